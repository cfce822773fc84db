/*
 * srmg: MI355X-native (gfx950) block-structured AMR geometric multigrid.
 *
 * C ABI of libsrmg.so. It replaces the reference's elliptic-solver entry
 * points, which a Julia caller reaches through ccall with plain pointers:
 *
 *   vcycle!(u, b, r, u1, s, tree, conn, geometry, bc, lpldisc;
 *           lmax=0, ndown=3, ntop=5, nup=3, omega=1.0)     src/multigrid.jl:60-106
 *   fmg!(u, b, r, u1, s, tree, conn, geometry, bc, lpldisc;
 *        lmax=0, ndown=3, ntop=5, nup=3, omega=1.25)        src/multigrid.jl:17-51
 *   resnorm!(r, b, u, s, tree, conn, geometry, bc, lpldisc)  src/multigrid.jl:391-417
 *   fill_ghost!(u, l, conn, bc)                              src/connectivity.jl:155-159
 *
 * The tree and connectivity stay on the host (src/tree.jl, src/connectivity.jl);
 * srmg_set_tree receives them at every regrid as the reference's isbits records
 * and rebuilds the device tables. Fields stay resident in HBM between calls.
 *
 * Conventions: every function returns 0 on success and a nonzero status on
 * failure, with a message in srmg_last_error() (the Julia wrapper turns it into
 * error(...), as the reference's asserts do). The host owns host buffers, the
 * library owns device buffers; no host pointer is retained after a call.
 * A context is single-caller and owns one HIP stream; calls are asynchronous
 * on that stream except where a host result is returned (resnorm, download).
 */
#ifndef SRMG_H
#define SRMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct srmg_ctx srmg_ctx;

/* field element types */
enum { SRMG_F64 = 0, SRMG_F32 = 1 };
/* geometry: CartesianGeometry, CylindricalGeometry{1}, CylindricalGeometry{2}
 * (src/geometry.jl:14-20) */
enum { SRMG_CARTESIAN = 0, SRMG_CYL1 = 1, SRMG_CYL2 = 2 };
/* discretization: LaplacianDiscretization{D,2}, HelmholtzDiscretization{D,2,T}
 * (src/laplacian.jl:14, src/helmholtz.jl:9-11) */
/* SRMG_LAPLACIAN4: the compact 4th-order LaplacianDiscretization{D,4}.
 * srmg_rhs applies its R (Cartesian and cylindrical); srmg_vcycle runs its
 * V-cycle with the FourColors smoother (src/multigrid.jl:206-232, 268-295) for
 * Cartesian and cylindrical geometry, FP64 fields, the reference's level-1
 * sweeps and single trees that are properly nested for the BoxStencil (its
 * ghosts: srmg_fill_ghost4), srmg_fmg its FMG, srmg_solve and
 * srmg_solve_mixed its solves to a tolerance (guess NONE or ZERO; the mixed
 * form with FP32 order-4 correction cycles), and srmg_resnorm its residual norm (r receives
 * the residual, as resnorm! overwrites r); the other solve entry points
 * refuse it. */
enum { SRMG_LAPLACIAN = 0, SRMG_HELMHOLTZ = 1, SRMG_LAPLACIAN4 = 2 };
/* coarsest-level treatment: the reference's ntop Gauss-Seidel sweeps
 * (src/multigrid.jl:89), or an exact FP64 solve of the level-1 problem
 * (an extension with no reference counterpart, needed when level 1 is large). */
enum { SRMG_COARSE_GS = 0, SRMG_COARSE_EXACT = 1 };

typedef struct {
    int32_t geometry;  /* SRMG_CARTESIAN / SRMG_CYL1 / SRMG_CYL2 */
    int32_t disc;      /* SRMG_LAPLACIAN / SRMG_HELMHOLTZ */
    double k2;         /* HelmholtzDiscretization.k2 (ignored for the Laplacian) */
    int8_t bc[27];     /* GeneralHomogeneousBoundaryConditions{D}.a (src/boundary.jl:12-14),
                          3^D Int8 table in Julia column-major order */
    int32_t coarse;    /* SRMG_COARSE_GS / SRMG_COARSE_EXACT */
} srmg_op;

/* Message for the last failing call on this thread. */
const char *srmg_last_error(void);

/* Context for D in {2,3} and block side M in {8,16} (ScalarBlockField{D,M,G,T},
 * src/blockfield.jl:16-74). Uses the current HIP device. */
int srmg_create(srmg_ctx **ctx, int D, int M);
int srmg_destroy(srmg_ctx *ctx);

/* Regrid (after connectivity!, src/streamer.jl:548). Records are concatenated
 * over levels; *_off has nlevels+1 entries (record offsets per level):
 *   pairs       (coord[D], blk)                      BlockLayer.pairs, src/blocklayer.jl:23
 *   neighbor    (from, to, face[D])                  Neighbor{D}, src/connectivity.jl:14-20
 *   boundary    (block, face[D], bnd[D], level)      Boundary{D}, src/connectivity.jl:6-11
 *   refboundary (fine, coarse, face[D], subblock[D]) RefBoundary{D}, src/connectivity.jl:23-32
 *   child       (fine, coarse, subblock[D])          Child{D}, src/connectivity.jl:36-42
 * Block ids are the reference's 1-based ids, at most nblocks. Level 1 must be
 * complete (populate!, src/tree.jl:31-44). Existing fields are reallocated
 * (contents lost): upload them again after a regrid. Records are checked
 * before the context changes: a refused call leaves the old tree and fields. */
int srmg_set_tree(srmg_ctx *ctx, int nlevels, int64_t nblocks,
                  const int64_t *pairs, const int64_t *pairs_off,
                  const int64_t *neighbor, const int64_t *neighbor_off,
                  const int64_t *boundary, const int64_t *boundary_off,
                  const int64_t *refboundary, const int64_t *refboundary_off,
                  const int64_t *child, const int64_t *child_off);

/* Several independent trees in one context (a forest): each tree is an
 * independent problem (own tree, connectivity and later its own s and
 * operator), and every level-wide launch covers the blocks of all trees, so
 * one launch per smoother step serves them all. Arrays have one entry per
 * tree; the records are as in srmg_set_tree, with each tree's own 1-based
 * block ids. Host fields of a forest are the trees' host arrays concatenated
 * in tree order. Level 1 of every tree must be complete. */
int srmg_set_forest(srmg_ctx *ctx, int ntrees, const int *nlevels, const int64_t *nblocks,
                    const int64_t *const *pairs, const int64_t *const *pairs_off,
                    const int64_t *const *neighbor, const int64_t *const *neighbor_off,
                    const int64_t *const *boundary, const int64_t *const *boundary_off,
                    const int64_t *const *refboundary, const int64_t *const *refboundary_off,
                    const int64_t *const *child, const int64_t *const *child_off);
/* Number of trees of the context. */
int srmg_tree_count(srmg_ctx *ctx, int *ntrees);

/* connectivity!(conn, tree, stencil) (src/connectivity.jl:106-149) on the
 * GPU: the four record lists of a tree from its blocks alone, in the
 * reference's order (level by level, blocks in `pairs` order, offsets in
 * stencil order), with the record layouts of srmg_set_tree. stencil:
 * SRMG_STAR_STENCIL (StarStencil{D}, the order-2 operators) or
 * SRMG_BOX_STENCIL (BoxStencil{D}, diagonals included). rootsize: the level-1
 * domain in blocks (level l spans rootsize * 2^(l-1)). The caller provides
 * each list with a capacity in records (at most nblocks * (number of stencil
 * offsets) for the first three, the blocks above level 1 for child) and its
 * nlevels+1 offsets. Fails on a position holding two blocks, a neighbour
 * position without a parent block (proper nesting broken) or a block without
 * a parent, as the reference asserts. Uses the current HIP device and a
 * stream of its own; returns when the records are on the host. */
enum { SRMG_STAR_STENCIL = 0, SRMG_BOX_STENCIL = 1 };
int srmg_connectivity(int D, int nlevels, const int64_t *rootsize,
                      const int64_t *pairs, const int64_t *pairs_off, int stencil,
                      int64_t *neighbor, int64_t *neighbor_off, int64_t neighbor_cap,
                      int64_t *boundary, int64_t *boundary_off, int64_t boundary_cap,
                      int64_t *refboundary, int64_t *refboundary_off, int64_t refboundary_cap,
                      int64_t *child, int64_t *child_off, int64_t child_cap);

/* fill_ghost!(u, l, conn, bc) for every level, ascending, for the order-4
 * operator: its depth-1 ghosts, corners and edges included, are kept
 * explicitly (copies from same-level neighbours, then boundary mirrors, then
 * coarse-fine interpolations, as the reference fills them with BoxStencil
 * records). Call it where the reference pre-fills u before the cycles
 * (src/streamer.jl:323-325). FP64 fields, single-tree contexts. */
int srmg_fill_ghost4(srmg_ctx *ctx, int field, const int8_t *bc);

/* restrict_full!(u, conn) (src/restrict.jl:43-47): every level's blocks
 * restricted into their parents, finest level first. */
int srmg_restrict_full(srmg_ctx *ctx, int field);

/* The right-hand side of derivs!'s Poisson block (src/streamer.jl:313-318):
 * for every level, fill_ghost!(b, l) with op->bc, then rhs_level!(b1, b, l)
 * (src/multigrid.jl:363-388): b1 = b for the order-2 operators, the compact
 * 4th-order R b for SRMG_LAPLACIAN4 (Cartesian 2D/3D, cylindrical 2D). b1's
 * ghosts are derived when read, as for every field. */
int srmg_rhs(srmg_ctx *ctx, int b1, int b, const srmg_op *op);

/* Regrid that carries fields across on the device: the field part of
 * applydelta!(tree, refdelta, freeblocks, fields, minlevel, maxlevel)
 * (src/refinement.jl:339-370, with interp! and restrict! of
 * src/streamer.jl:191-221) around srmg_set_tree. The records describe the tree
 * after the host's tree edit and connectivity! (as for srmg_set_tree); blocks
 * are matched to the old tree by (level, coordinate), so block ids may be
 * reused. For each listed field: blocks that keep their (level, coordinate)
 * keep their values; with SRMG_REGRID_INTERP a new block is interp!'d from its
 * parent's old values and depth-1 ghosts (filled first as fill_ghost! does,
 * with bc), otherwise it starts at zero; with SRMG_REGRID_RESTRICT a removed
 * block is restrict!'ed into its parent's subblock. A new block's parent must
 * be in the old tree and a removed block's parent must be kept (one level per
 * regrid, as applydelta! does); this, the block coordinates and every check
 * of srmg_set_tree are made before anything changes: a refused regrid leaves
 * the context with its old tree and field values.
 * Fields not listed are reallocated as in srmg_set_tree. bc may be NULL
 * when no field is interpolated. */
enum { SRMG_REGRID_INTERP = 1, SRMG_REGRID_RESTRICT = 2 };
int srmg_regrid(srmg_ctx *ctx, int nlevels, int64_t nblocks,
                const int64_t *pairs, const int64_t *pairs_off,
                const int64_t *neighbor, const int64_t *neighbor_off,
                const int64_t *boundary, const int64_t *boundary_off,
                const int64_t *refboundary, const int64_t *refboundary_off,
                const int64_t *child, const int64_t *child_off,
                const int *fields, const int *modes, int nfields, const int8_t *bc);
/* srmg_regrid for a forest (srmg_set_forest's arguments): tree t of the new
 * forest replaces tree t of the context, each matched by (level, coordinate);
 * the number of trees stays. */
int srmg_regrid_forest(srmg_ctx *ctx, int ntrees, const int *nlevels, const int64_t *nblocks,
                       const int64_t *const *pairs, const int64_t *const *pairs_off,
                       const int64_t *const *neighbor, const int64_t *const *neighbor_off,
                       const int64_t *const *boundary, const int64_t *const *boundary_off,
                       const int64_t *const *refboundary, const int64_t *const *refboundary_off,
                       const int64_t *const *child, const int64_t *const *child_off,
                       const int *fields, const int *modes, int nfields, const int8_t *bc);

/* Device-resident block field of the given element type; returns its id. */
int srmg_field_create(srmg_ctx *ctx, int dtype, int *id);
int srmg_field_destroy(srmg_ctx *ctx, int id);

/* Host <-> device copies in the reference's host layout: column-major
 * S^D x nblocks with S = M + 2G, block blk at offset (blk-1) S^D
 * (src/blockfield.jl:33-59). Upload reads valid cells and the depth-1 ghost
 * layer of every coarse-fine face; download writes valid cells and the depth-1
 * ghost layer of every face (same-level copy, boundary mirror, coarse-fine
 * interpolation), leaving deeper ghost layers untouched. host_dtype may differ
 * from the field's type (converted). */
int srmg_upload(srmg_ctx *ctx, int id, const void *host, int host_dtype, int G);
int srmg_download(srmg_ctx *ctx, int id, void *host, int host_dtype, int G);
/* As srmg_download, and also the depth-1 ghosts of domain-boundary faces,
 * from the boundary table bc (3^D Int8, as in srmg_op). */
int srmg_download_bc(srmg_ctx *ctx, int id, void *host, int host_dtype, int G, const int8_t *bc);
/* Set a field to zero (valid cells and ghost buffers). */
int srmg_zero(srmg_ctx *ctx, int id);

/* fill_ghost!(u, level, conn, bc) (src/connectivity.jl:155-159). Same-level and
 * boundary ghosts are always derived from the current valid cells; this call
 * refreshes the coarse-fine interpolation ghosts of `level` from level-1. */
int srmg_fill_ghost(srmg_ctx *ctx, int id, int level, const int8_t *bc);
/* fill_ghost! for l = 1..nlevels, the callers' pre-solve fill
 * (src/streamer.jl:323-325, src/photo.jl:49-51). */
int srmg_fill_ghost_all(srmg_ctx *ctx, int id, const int8_t *bc);

/* vcycle! (src/multigrid.jl:60-106): L u + s b = 0. Mutates u, b (the FAS
 * right-hand side at covered coarse cells), r and u1 (scratch). lmax = 0 means
 * all levels. All four fields must have the same element type. */
int srmg_vcycle(srmg_ctx *ctx, int u, int b, int r, int u1, double s, const srmg_op *op,
                int lmax, int ndown, int ntop, int nup, double omega);

/* vcycle! on every tree of a forest at once, tree t with scale s[t] and
 * operator ops[t] (geometry, boundary table and coarse treatment must agree;
 * k2 and the discretization may differ, e.g. the terms of a photoionization
 * model). Each tree gets exactly its own vcycle!. */
int srmg_vcycle_batched(srmg_ctx *ctx, int u, int b, int r, int u1, const double *s, const srmg_op *ops,
                        int lmax, int ndown, int ntop, int nup, double omega);

/* fmg! (src/multigrid.jl:17-51). */
int srmg_fmg(srmg_ctx *ctx, int u, int b, int r, int u1, double s, const srmg_op *op,
             int ndown, int ntop, int nup, double omega);

/* resnorm! (src/multigrid.jl:391-417): sqrt of the sum over leaf cells of
 * r^2 2^(-D(l-1)), r = -(b + L u / s_l). Synchronizes. Does not modify r
 * (the reference overwrites its r argument with squared residuals). */
int srmg_resnorm(srmg_ctx *ctx, int r, int b, int u, double s, const srmg_op *op, double *out);
/* resnorm! of every tree of a forest: out[t]. */
int srmg_resnorm_batched(srmg_ctx *ctx, int r, int b, int u, const double *s, const srmg_op *ops, double *out);

/* Block until all work queued on the context's stream has finished. */
int srmg_synchronize(srmg_ctx *ctx);

/* ---- solve drivers (the callers' loops of src/streamer.jl:322-334) ---- */

typedef struct {
    int32_t ndown, ntop, nup;  /* (2, 2, 4) in src/simulate.jl:145 order (nup, ndown, ntop) */
    double omega;              /* 1.0 for V-cycles */
    double tol;                /* stop when resnorm(u) <= tol * resnorm(u = 0) */
    int32_t max_cycles;
    int32_t use_graph;         /* capture one cycle in a hipGraph and replay it */
    int32_t inner_cycles;      /* mixed precision: FP32 V-cycles per FP64 correction (0 means 1) */
    int32_t guess;             /* first iterate: SRMG_GUESS_* */
    int32_t guess_smooth;      /* Gauss-Seidel sweeps on level 1 after a learned first iterate */
    int32_t first_fmg;         /* 1: the first cycle is fmg! (src/multigrid.jl:17-51) from the first
                                  iterate, the later ones vcycle!; in the mixed form the first
                                  correction runs fmg! on e before its inner_cycles V-cycles.
                                  Order-2 operators. */
    int32_t inner_first;       /* mixed precision: FP32 V-cycles of the first correction (0 means
                                  inner_cycles); the first correction starts from the largest
                                  defect, where the FP32 floor stops progress sooner */
} srmg_solve_opts;

/* First iterate of a solve. NONE: u as given. ZERO: u = 0 whatever u holds
 * (its residual is -b, so the initial norm and FP64 defect come from b alone).
 * LEARNED: the network attached
 * with srmg_set_guess predicts the level-1 solution from the level-1 right-hand
 * side; it is interpolated to every finer level. COARSE: the same with the
 * exact level-1 solve in place of the network. A learned or coarse first
 * iterate replaces u; a tree whose first-iterate residual is not finite or
 * exceeds resnorm(u = 0) starts from u = 0 instead (counted by
 * srmg_guess_fallbacks). No counterpart in the reference (its solves start
 * from the previous u). */
enum { SRMG_GUESS_NONE = 0, SRMG_GUESS_LEARNED = 1, SRMG_GUESS_COARSE = 2, SRMG_GUESS_ZERO = 3 };

/* Solve from the current u (ghosts pre-filled here) with plain V-cycles of
 * the fields' own type until the relative residual is <= tol. hist (may be
 * NULL, else max_cycles+1 entries) receives resnorm/resnorm0 after each cycle
 * (hist[0] = resnorm(u_initial)/resnorm0). */
int srmg_solve(srmg_ctx *ctx, int u, int b, int r, int u1, double s, const srmg_op *op,
               const srmg_solve_opts *opts, int *cycles, double *hist);

/* srmg_solve on every tree of a forest (per-tree s and operator) until all
 * trees reach tol; hist (may be NULL) has ntrees x (max_cycles+1) entries. */
int srmg_solve_batched(srmg_ctx *ctx, int u, int b, int r, int u1, const double *s, const srmg_op *ops,
                       const srmg_solve_opts *opts, int *cycles, double *hist);

/* Mixed precision: u, b FP64; e, be, re, ue FP32 scratch fields. Each outer
 * step forms the FP64 residual, runs inner_cycles FP32 V-cycles (inner_first in
 * the first step, when set) on the
 * correction equation from e = 0, adds e to u and restricts u so covered
 * cells equal the mean of their children; the coarsest-level solve is FP64.
 * cycles counts outer steps; hist has one entry per outer step. */
int srmg_solve_mixed(srmg_ctx *ctx, int u, int b, int e, int be, int re, int ue, double s,
                     const srmg_op *op, const srmg_solve_opts *opts, int *cycles, double *hist);
/* srmg_solve_mixed on every tree of a forest; hist as in srmg_solve_batched. */
int srmg_solve_mixed_batched(srmg_ctx *ctx, int u, int b, int e, int be, int re, int ue, const double *s,
                             const srmg_op *ops, const srmg_solve_opts *opts, int *cycles, double *hist);

/* ---- electric field (src/multigrid.jl:421-451, caller src/streamer.jl:340-351) ---- */

/* E = -grad u + e0 on the cell faces and |E| at the cell centres of every block
 * of every level, for a constant external field e0[D] and root cell size h
 * (level l uses h / 2^(l-1)). u (FP64) must hold current coarse-fine ghost
 * slabs (true after any solve or upload); bc gives its boundary ghosts (the
 * Poisson bc). |E| goes to the FP64 field eabs (valid cells). If e_host is not
 * NULL, E is written to it in the reference's VectorBlockField{D,M,G} layout
 * ((M+2G+1)^D x D x nblocks, column-major; only the valid faces
 * G+1..G+M+1 along the component's dimension are written). */
int srmg_electric_field(srmg_ctx *ctx, int u, int eabs, double h, const double *e0, const int8_t *bc,
                        double *e_host, int G);

/* ---- free boundary conditions (src/freebc.jl; 2D cylindrical, caller
 *      src/streamer.jl:320 and 340-349) ---- */

typedef struct srmg_freebc srmg_freebc;

/* FreeBC constructor (src/freebc.jl:32-54): dim is the radial dimension of
 * CylindricalGeometry{dim} (1 or 2), level the level at which the boundary is
 * resolved (refine_maxlevel), rootsize[2] the level-1 block domain, h the
 * level-1 cell size and f[n] the boundary kernel
 * f_k = (2/H) / (k (I1/I0 + K1/K0)(k R)), k = pi m / H, m = 1..n, with
 * n = rootsize[other] 2^(level-1) M (computed by the caller, as the reference
 * computes it with SpecialFunctions). */
int srmg_freebc_create(srmg_ctx *ctx, srmg_freebc **fb, int dim, int level, const int64_t *rootsize, double h,
                       const double *f, int64_t n);
int srmg_freebc_destroy(srmg_freebc *fb);
/* freebc!(q, u, h, s, tree, conn, fb) (src/freebc.jl:87-110): the outer
 * values of u over h_l (getouter!), the two sine transforms with the kernel f,
 * the boundary values b, and setouter! on the FP64 field q (outermost cells and
 * the corner cells at coarse-fine faces). u is the homogeneous solution. */
int srmg_freebc_apply(srmg_ctx *ctx, srmg_freebc *fb, int q, int u, double s);
/* The boundary values b[n] of the last srmg_freebc_apply. */
int srmg_freebc_boundary(srmg_freebc *fb, double *b, int64_t n);
/* srmg_electric_field with the ghost fill of fill_ghost_free! (src/freebc.jl:218-238):
 * the ghosts beyond the outer radial face gain 2 mean(b) before E is formed. */
int srmg_freebc_electric_field(srmg_ctx *ctx, srmg_freebc *fb, int u, int eabs, double h, const double *e0,
                               const int8_t *bc, double *e_host, int G);

/* ---- learned first iterate ---- */

/* Attach (net != NULL) or detach the network used by SRMG_GUESS_LEARNED. The
 * context does not own it; it must outlive its use. */
typedef struct srmg_unet srmg_unet;
int srmg_set_guess(srmg_ctx *ctx, srmg_unet *net);
/* Trees of the last solve whose learned/coarse first iterate was rejected. */
int srmg_guess_fallbacks(srmg_ctx *ctx, int *n);

/* Exact solve of L_1 u = f on ngrid stacked uniform level-1 grids (x fastest,
 * n[0] x n[1] (x n[2]) cells, the level-1 stencil of op with scale s: the
 * Helmholtz shift is -s k2), FP64 device pointers. Training targets for the
 * network; f = -s b for the problem of srmg_solve. */
int srmg_grid_solve(srmg_ctx *ctx, int D, const int *n, double s, const srmg_op *op, const double *f, double *u,
                    int ngrid);

/* srmg_grid_solve without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * and training-data loops that forbid hipMalloc). The transform operands and
 * the intermediate grids live in a file-scope device array of 16 MiB, one per
 * GPU, held until the process ends; the call fails (nothing launched) when
 * the grids need more: sum over axes of (2 n_d^2 + n_d) + ngrid (1 + 2 cells)
 * doubles, each piece rounded up to 256 bytes. Same results as
 * srmg_grid_solve. Every call on one GPU shares that array: calls that
 * overlap on one GPU (other threads, other contexts) race. */
int srmg_grid_solve_noalloc(srmg_ctx *ctx, int D, const int *n, double s, const srmg_op *op, const double *f,
                            double *u, int ngrid);

/* U-Net of 3^D convolutions on a D-dimensional grid n[0] x n[1] (x n[2]):
 * `levels` resolution levels with widths[l] channels (powers of two in 8..64;
 * extents divisible by 2^(levels-1)), BF16 matrix-core kernels with FP32
 * accumulation, FP32 master parameters initialized from `seed`. */
int srmg_unet_create(srmg_unet **net, int D, const int *n, int levels, const int *widths, uint64_t seed);
int srmg_unet_destroy(srmg_unet *net);
int srmg_unet_param_count(const srmg_unet *net, int64_t *count);
/* FP32 parameters: per convolution W[cout][tap][cin] then b[cout], taps
 * t = (dx+1) + 3 (dy+1) + 9 (dz+1); convolutions in the order encoder levels
 * 0..L-1 (two each), decoder levels L-2..0 (coarse linear, then ReLU), output. */
int srmg_unet_get_params(srmg_unet *net, float *host);
int srmg_unet_set_params(srmg_unet *net, const float *host);  /* also resets the optimizer */
/* Output scale c_out: the network predicts u / (max|f| c_out). */
int srmg_unet_set_output_scale(srmg_unet *net, double c_out);
int srmg_unet_get_output_scale(const srmg_unet *net, double *c_out);
/* The level-1 operator the network is trained for (geometry, disc, k2, bc);
 * a learned first iterate is refused for any other. */
int srmg_unet_set_problem(srmg_unet *net, const srmg_op *op);
/* u = prediction for `batch` stacked grids f (FP32 device pointers). */
int srmg_unet_forward(srmg_unet *net, const float *f, float *u, int batch);
/* Mean squared error of the normalized prediction against the targets u and
 * its gradient (host, param_count entries; either output may be NULL). */
int srmg_unet_loss_grad(srmg_unet *net, const float *f, const float *u, int batch, double *loss, float *grad);
/* One Adam step (beta 0.9, 0.999) on the batch; loss (may be NULL) before the step. */
int srmg_unet_train_step(srmg_unet *net, const float *f, const float *u, int batch, float lr, double *loss);
int srmg_unet_synchronize(srmg_unet *net);

/* ---- measurement ---- */

enum { SRMG_KERNEL_GS_PASS = 0, SRMG_KERNEL_RESIDUAL = 1, SRMG_KERNEL_GS_SWEEP = 2, SRMG_KERNEL_STREAM = 3,
       SRMG_KERNEL_RESIDUAL_GATHER = 4, SRMG_KERNEL_STREAM16 = 5 };

/* Average duration of one launch of a level kernel, timed with HIP events on
 * the context's stream over `reps` back-to-back launches (for the fused sweep
 * and the stream kernels: captured in a graph and replayed, as solves run
 * them, the faster of two replays): one colour of the
 * Gauss-Seidel sweep (alternating), the residual (LDS-tiled, or the per-cell
 * gather form), one fused red-black sweep, or the stream kernel r = u + b
 * over the level (bandwidth calibration; STREAM16: 16-byte loads and
 * nontemporal 16-byte stores per lane).
 * Modifies u (smoothers) or r (residual, stream). */
int srmg_time_kernel(srmg_ctx *ctx, int kind, int level, int u, int b, int r, double s,
                     const srmg_op *op, int reps, double *ms_per_launch);

/* Autotuner: times every compiled tile shape of the fused sweep on every
 * level with `reps` sweeps of u (captured in a graph and replayed, as solves
 * run them), keeps the fastest per level for later solves
 * of fields of u's element type, and writes the choices to chosen[nlevels] if
 * non-NULL. u is restored. Shapes: 0..3 blocks per work group x work-group
 * size; 4, 5, 6 (3D M=8 and 2D M=16) one sibling group of 2^D blocks as one
 * tile on 256 / 512 / 1024 threads; 7 (3D M=8) one sibling group on 256
 * threads, one thread per column of the group; 8 (3D M=8) the same on 512
 * threads, two per column; 9, 10, 11 (3D M=8) shape 8 with nontemporal
 * stores of the result, nontemporal loads of the right-hand side, or both
 * (levels without complete sibling groups, and thread counts that do not
 * divide a group's cells per colour, run 0; 7 to 11 run 4 and 5 where the
 * column forms do not apply). */
int srmg_autotune(srmg_ctx *ctx, int u, int b, double s, const srmg_op *op, int reps, int *chosen);
/* Force the tile shape (0..11) of the fused sweep for one element type and level (1-based).
 * Until autotuned or forced, 3D M=8 contexts run shape 8 and the others shape 0. */
int srmg_set_sweep_config(srmg_ctx *ctx, int dtype, int level, int cfg);
/* 1 when the level's blocks form complete sibling groups (tile shapes 4..8 apply). */
int srmg_level_grouped(srmg_ctx *ctx, int level, int *grouped);
/* Use (default) or bypass the sibling-group forms of the sweep, residual and
 * prolongation kernels on grouped levels; both give the same bits. */
int srmg_set_group_tiles(srmg_ctx *ctx, int enable);
/* Kernels of the level-1 solve (exact coarsest solve and srmg_grid_solve):
 * 2 (default) three launches for 3D grids whose extents are multiples of 8 and
 * at most 64 (plane and column tiles in LDS: forward x and y, forward z with
 * the eigenvalue division and inverse z, inverse y and x), else as 1; 1 one
 * launch per axis transform, one wave per work group without LDS where the
 * axis length is at most 64; 0 the LDS-tiled axis kernels. All give the same
 * bits. */
int srmg_set_axis_gemm(srmg_ctx *ctx, int variant);
/* Order in which the column sweeps (tile shapes 8..11) of grouped 3D levels
 * visit the sibling groups: 0 (default) the device block order (Morton); 1
 * slabs two groups thick in z, row by row, so a group's x and y neighbours
 * are read close in time (less HBM traffic, not faster on cfg4). Both give
 * the same bits. No counterpart in the reference (a tuning switch like the
 * two above). */
int srmg_set_group_order(srmg_ctx *ctx, int order);

/* Blocks, leaf blocks and coarse-fine ghost slabs of a level (1-based). */
int srmg_level_info(srmg_ctx *ctx, int level, int64_t *nblocks, int64_t *nleaf, int64_t *nslots);

/* ---- host utilities (no device work) ---- */

/* Eigendecomposition of the level-1 1D operator along one axis with n cells
 * (radial != 0: cylindrical weights), boundary signs alow/ahigh: T = P diag(lam) Q,
 * P and Q row-major n x n. Used by the exact coarsest-level solve. */
int srmg_axis_decomp(int n, int radial, double alow, double ahigh,
                     double *lam, double *P, double *Q);

#ifdef __cplusplus
}
#endif

#endif /* SRMG_H */
