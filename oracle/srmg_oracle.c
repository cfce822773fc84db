/*
 * srmg oracle: a plain-C restatement of the reference's block-structured AMR
 * FAS multigrid (Poisson and Helmholtz, order 2), operating on the reference's
 * own host memory layout.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library: only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it. The
 * product path never calls it.
 *
 * Layout: a field is the reference's ScalarBlockField{D,M,G,Float64} with
 * contiguous storage (src/blockfield.jl:33-59): column-major S^D x nblk,
 * S = M + 2G, first index fastest, block ids 1-based. All local indices below
 * are 1-based, as in the reference.
 *
 * Tree and connectivity arrive as the reference's isbits records (int64):
 *   pairs       (coord[D], blk)                      src/blocklayer.jl:23
 *   Neighbor    (from, to, face[D])                  src/connectivity.jl:14-20
 *   Boundary    (block, face[D], bnd[D], level)      src/connectivity.jl:6-11
 *   RefBoundary (fine, coarse, face[D], subblock[D]) src/connectivity.jl:23-32
 *   Child       (fine, coarse, subblock[D])          src/connectivity.jl:36-42
 * each concatenated over levels with an (nlevels + 1) offset array.
 *
 * Floating point: compiled with -ffp-contract=off; fma() appears exactly where
 * the reference writes muladd (src/multigrid.jl:176-177). Stencil terms are
 * summed in the order the reference's @stencil macro emits them
 * (src/laplacian.jl:25-49: column-major over the 3^D literal, zeros skipped).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int D, M, G, S, nlevels;
    int64_t SD;        /* S^D */
    int64_t nblocks;   /* storage blocks */
    int64_t *pairs, *pairs_off;
    int64_t *nbr, *nbr_off;
    int64_t *bnd, *bnd_off;
    int64_t *ref, *ref_off;
    int64_t *chd, *chd_off;
    int64_t *coord_of; /* blk -> coord[D] (index (blk-1)*D) */
    int *level_of;     /* blk -> level */
    /* cached level-1 axis decompositions, keyed by (geom, bc) */
    int ck_valid, ck_geom;
    int8_t ck_bc[27];
    double *ck_lam[3], *ck_P[3], *ck_Q[3];
} ora_tree;

typedef struct {
    int geom;          /* 0 Cartesian; 1 CylindricalGeometry{1}; 2 CylindricalGeometry{2} */
    int helm;          /* 0 LaplacianDiscretization{D,2}; 1 HelmholtzDiscretization{D,2} */
    double k2;
    int8_t bc[27];     /* GeneralHomogeneousBoundaryConditions{D}.a, column-major */
    int order;         /* 4: LaplacianDiscretization{D,4} (Cartesian); otherwise order 2 */
} ora_op;

/* OpenMP threads of the level loops (the CPU baseline times 1 and all) */
void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int ora_max_threads(void) { return omp_get_max_threads(); }

static int64_t *dup64(const int64_t *src, int64_t n) {
    int64_t *p = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (n > 0) memcpy(p, src, (size_t)n * sizeof(int64_t));
    return p;
}

ora_tree *ora_tree_create(int D, int M, int G, int nlevels, int64_t nblocks,
                          const int64_t *pairs, const int64_t *pairs_off,
                          const int64_t *nbr, const int64_t *nbr_off,
                          const int64_t *bnd, const int64_t *bnd_off,
                          const int64_t *ref, const int64_t *ref_off,
                          const int64_t *chd, const int64_t *chd_off) {
    ora_tree *t = (ora_tree *)calloc(1, sizeof(ora_tree));
    t->D = D; t->M = M; t->G = G; t->S = M + 2 * G; t->nlevels = nlevels;
    t->SD = 1;
    for (int d = 0; d < D; d++) t->SD *= t->S;
    t->nblocks = nblocks;
    t->pairs_off = dup64(pairs_off, nlevels + 1);
    t->nbr_off = dup64(nbr_off, nlevels + 1);
    t->bnd_off = dup64(bnd_off, nlevels + 1);
    t->ref_off = dup64(ref_off, nlevels + 1);
    t->chd_off = dup64(chd_off, nlevels + 1);
    t->pairs = dup64(pairs, pairs_off[nlevels] * (D + 1));
    t->nbr = dup64(nbr, nbr_off[nlevels] * (2 + D));
    t->bnd = dup64(bnd, bnd_off[nlevels] * (2 + 2 * D));
    t->ref = dup64(ref, ref_off[nlevels] * (2 + 2 * D));
    t->chd = dup64(chd, chd_off[nlevels] * (2 + D));
    t->coord_of = (int64_t *)calloc((size_t)(nblocks * D + 1), sizeof(int64_t));
    t->level_of = (int *)calloc((size_t)(nblocks + 1), sizeof(int));
    for (int l = 0; l < nlevels; l++) {
        for (int64_t k = pairs_off[l]; k < pairs_off[l + 1]; k++) {
            const int64_t *p = t->pairs + k * (D + 1);
            int64_t blk = p[D];
            for (int d = 0; d < D; d++) t->coord_of[(blk - 1) * D + d] = p[d];
            t->level_of[blk - 1] = l + 1;
        }
    }
    return t;
}

void ora_tree_destroy(ora_tree *t) {
    if (!t) return;
    free(t->pairs); free(t->pairs_off); free(t->nbr); free(t->nbr_off);
    free(t->bnd); free(t->bnd_off); free(t->ref); free(t->ref_off);
    free(t->chd); free(t->chd_off); free(t->coord_of); free(t->level_of);
    for (int d = 0; d < 3; d++) { free(t->ck_lam[d]); free(t->ck_P[d]); free(t->ck_Q[d]); }
    free(t);
}

/* offset of local 1-based index I (length D) in block blk (1-based) */
static inline int64_t OFF(const ora_tree *t, const int *I, int64_t blk) {
    int64_t o = (I[0] - 1);
    if (t->D >= 2) o += (int64_t)(I[1] - 1) * t->S;
    if (t->D >= 3) o += (int64_t)(I[2] - 1) * t->S * t->S;
    return o + (blk - 1) * t->SD;
}

/* ---------------------------------------------------------------------- */
/* Operators: src/laplacian.jl:78-128, 278-303; src/helmholtz.jl:17-100      */
/* ---------------------------------------------------------------------- */

/* Centre coefficient of the stencil / diagonal element (src/laplacian.jl:94,
 * 127, 302; src/helmholtz.jl:34-35, 66-67, 99-100). */
static inline double centre_coef(const ora_tree *t, const ora_op *op, double s) {
    double c = -2.0 * t->D;
    if (op->helm) c = c - s * op->k2;
    return c;
}

/* applystencil(u[blk], s, I, J, ld, geom, Val(:lhs)). The @stencil macro emits
 * v = 0; v += a[S1] * u[I + S1 - 2] for S1 in CartesianIndices(a) (column-major)
 * with a[S1] != 0. For the 5-point 2D literal this is the offset order
 * (0,-1), (-1,0), (0,0), (+1,0), (0,+1); for the 7-point 3D literal
 * (0,0,-1), (0,-1,0), (-1,0,0), (0,0,0), (+1,0,0), (0,+1,0), (0,0,+1). */
static double stencil_sum(const ora_tree *t, const ora_op *op, const double *u,
                          int64_t blk, const int *I, const int64_t *J, double cc);

static double stencil4(const ora_tree *t, const double *u, int64_t blk, const int *I);
static double stencil4cyl(const ora_tree *t, int N, const double *u, int64_t blk, const int *I, int64_t i);

static double applystencil(const ora_tree *t, const ora_op *op, const double *u,
                           int64_t blk, const int *I, const int64_t *J, double s) {
    if (op->order == 4 && op->geom) return stencil4cyl(t, op->geom - 1, u, blk, I, J[op->geom - 1]);
    if (op->order == 4) return stencil4(t, u, blk, I);
    return stencil_sum(t, op, u, blk, I, J, centre_coef(t, op, s));
}

/* applystencil for LaplacianDiscretization{D,4} x CartesianGeometry
 * (src/laplacian.jl:133-141, 305-320): the @stencil literal's nonzero terms in
 * column-major order (first offset fastest), then one division by 6. 2D:
 * [1 4 1; 4 -20 4; 1 4 1]; 3D: [0 1 0; 1 2 1; 0 1 0] in the outer slices and
 * [1 2 1; 2 -24 2; 1 2 1] in the middle one. */
static double stencil4(const ora_tree *t, const double *u, int64_t blk, const int *I) {
    const int D = t->D;
    double v = 0.0;
    int K[3] = {I[0], I[1], I[2]};
    const int nk = D == 3 ? 1 : 0;
    for (int k = -nk; k <= nk; k++)
        for (int j = -1; j <= 1; j++)
            for (int i = -1; i <= 1; i++) {
                const int n = (i != 0) + (j != 0) + (k != 0);
                double a;
                if (D == 2) a = n == 0 ? -20.0 : (n == 1 ? 4.0 : 1.0);
                else if (k == 0) a = n == 0 ? -24.0 : (n == 1 ? 2.0 : 1.0);
                else a = n == 1 ? 2.0 : (n == 2 ? 1.0 : 0.0);
                if (a == 0.0) continue;
                K[0] = I[0] + i;
                K[1] = I[1] + j;
                if (D == 3) K[2] = I[2] + k;
                v += a * u[OFF(t, K, blk)];
            }
    return v / 6.0;
}

/* diagelm of LaplacianDiscretization{D,4} x CartesianGeometry
 * (src/laplacian.jl:151-152, 339-340) */
static inline double diag4(int D) { return D == 2 ? -20.0 / 6.0 : -4.0; }

/* LaplacianDiscretization{2,4} x CylindricalGeometry{N+1} (src/laplacian.jl:156-200)
 * at 1-based global radial index i: the 3x3 literal, entry (a, b) at offset
 * (a - 1, b - 1) with a the first dimension, visited column-major; the
 * rational coefficients are integer ratios rounded once, the sum divided by 6 */
static void cyl4_coefs(int N, int64_t i, double w[3][3]) {
    const double lo = (double)(2 * (i - 1)) / (double)(2 * i - 1), hi = (double)(2 * i) / (double)(2 * i - 1);
    const int64_t q = 2 * i - 1;
    const double pm = (double)(4 * (-2 + 9 * i - 16 * i * i + 8 * i * i * i)) / (double)(q * q * q);
    const double pp = (double)(4 * (1 + i - 8 * i * i + 8 * i * i * i)) / (double)(q * q * q);
    const double cc = (double)(-16 * (1 - 5 * i + 5 * i * i)) / (double)((1 - 2 * i) * (1 - 2 * i));
    /* w[a][b]: a = offset + 1 in the first dimension, b in the second */
    if (N == 0) {  /* radial first dimension: rows are radial */
        w[0][0] = lo; w[0][1] = pm; w[0][2] = lo;
        w[1][0] = 4;  w[1][1] = cc; w[1][2] = 4;
        w[2][0] = hi; w[2][1] = pp; w[2][2] = hi;
    } else {       /* radial second dimension: columns are radial */
        w[0][0] = lo; w[0][1] = 4;  w[0][2] = hi;
        w[1][0] = pm; w[1][1] = cc; w[1][2] = pp;
        w[2][0] = lo; w[2][1] = 4;  w[2][2] = hi;
    }
}

static double stencil4cyl(const ora_tree *t, int N, const double *u, int64_t blk, const int *I, int64_t i) {
    double w[3][3];
    cyl4_coefs(N, i, w);
    double v = 0.0;
    int K[3] = {I[0], I[1], I[2]};
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
            K[0] = I[0] + a - 1;
            K[1] = I[1] + b - 1;
            v += w[a][b] * u[OFF(t, K, blk)];
        }
    return v / 6.0;
}

/* its diagelm: -16 (1 - 5i + 5i^2) / (1 - 2i)^2 / 6 (src/laplacian.jl:197-201) */
static inline double diag4cyl(int64_t i) {
    return (double)(-16 * (1 - 5 * i + 5 * i * i)) / (double)((1 - 2 * i) * (1 - 2 * i)) / 6.0;
}

/* the star-stencil sum with centre weight cc, in the @stencil order above */
static double stencil_sum(const ora_tree *t, const ora_op *op, const double *u,
                          int64_t blk, const int *I, const int64_t *J, double cc) {
    const int D = t->D;
    /* per-dimension (minus, plus) weights */
    double wm[3] = {1.0, 1.0, 1.0}, wp[3] = {1.0, 1.0, 1.0};
    if (op->geom == 1 || op->geom == 2) {
        int N = op->geom - 1;
        int64_t i = J[N];
        wm[N] = (double)(2 * i - 2) / (double)(2 * i - 1);
        wp[N] = (double)(2 * i) / (double)(2 * i - 1);
    }
    double v = 0.0;
    int K[3];
    /* minus offsets, highest dimension first */
    for (int d = D - 1; d >= 0; d--) {
        for (int e = 0; e < D; e++) K[e] = I[e];
        K[d] -= 1;
        v += wm[d] * u[OFF(t, K, blk)];
    }
    v += cc * u[OFF(t, I, blk)];
    for (int d = 0; d < D; d++) {
        for (int e = 0; e < D; e++) K[e] = I[e];
        K[d] += 1;
        v += wp[d] * u[OFF(t, K, blk)];
    }
    return v; /* divided by denom = 1 in the reference: exact */
}

/* J = I - G + gbl0 - 1 with gbl0 = M (b - 1) + 1 (src/multigrid.jl:170, src/mesh.jl:76-78) */
static inline void global_index(const ora_tree *t, int64_t blk, const int *I, int64_t *J) {
    for (int d = 0; d < t->D; d++)
        J[d] = I[d] - t->G + t->M * (t->coord_of[(blk - 1) * t->D + d] - 1);
}

/* ---------------------------------------------------------------------- */
/* Ghost filling: src/connectivity.jl:155-258, src/boundary.jl:45-47,       */
/* src/blockfield.jl:416-541                                                */
/* ---------------------------------------------------------------------- */

void ora_fill_ghost_copy(const ora_tree *t, double *u, int level) {
    const int D = t->D, M = t->M, G = t->G;
    const int W = 2 + D;
    int64_t lo = t->nbr_off[level - 1], hi = t->nbr_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        const int64_t *e = t->nbr + k * W;
        int64_t from = e[0], to = e[1];
        const int64_t *f = e + 2;
        int n[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) n[d] = (f[d] == 0) ? M : G;
        int i[3];
        for (i[2] = 1; i[2] <= n[2]; i[2]++)
        for (i[1] = 1; i[1] <= n[1]; i[1]++)
        for (i[0] = 1; i[0] <= n[0]; i[0]++) {
            int If[3], It[3];
            for (int d = 0; d < D; d++) {
                If[d] = (f[d] == 0) ? G + i[d] : (f[d] == -1 ? i[d] + G : i[d] + M);
                It[d] = (f[d] == 0) ? G + i[d] : (f[d] == 1 ? i[d] : i[d] + M + G);
            }
            u[OFF(t, It, to)] = u[OFF(t, If, from)];
        }
    }
}

static inline int bc_index(int D, const int64_t *bnd) {
    int idx = 0, m = 1;
    for (int d = 0; d < D; d++) { idx += (int)(bnd[d] + 1) * m; m *= 3; }
    return idx;
}

static inline int mirror_index(int G, int M, int i, int64_t dir) {
    if (dir == -1) return 2 * G - i + 1;
    if (dir == 1) return 2 * (G + M) - i + 1;
    return i;
}

void ora_fill_ghost_bnd(const ora_tree *t, double *u, int level, const int8_t *bc) {
    const int D = t->D, M = t->M, G = t->G;
    const int W = 2 + 2 * D;
    int64_t lo = t->bnd_off[level - 1], hi = t->bnd_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        const int64_t *e = t->bnd + k * W;
        int64_t blk = e[0];
        const int64_t *face = e + 1, *bnd = e + 1 + D;
        double a = (double)bc[bc_index(D, bnd)];
        int first[3] = {1, 1, 1}, last[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) {
            if (face[d] == 0) { first[d] = G + 1; last[d] = G + M; }
            else if (face[d] == -1) { first[d] = 1; last[d] = G; }
            else { first[d] = M + G + 1; last[d] = M + 2 * G; }
        }
        int g[3];
        for (g[2] = first[2]; g[2] <= last[2]; g[2]++)
        for (g[1] = first[1]; g[1] <= last[1]; g[1]++)
        for (g[0] = first[0]; g[0] <= last[0]; g[0]++) {
            int v[3];
            for (int d = 0; d < D; d++) v[d] = mirror_index(G, M, g[d], bnd[d]);
            u[OFF(t, g, blk)] = a * u[OFF(t, v, blk)];
        }
    }
}

/* interp!(::InterpLinear, dest, idest, src, isrc, func) (src/interp.jl:7-11, 102-141).
 * first_d/len_d describe the destination and source rectangles (actual 1-based
 * indices); add != 0 selects func = + (src/interp.jl:185-193). */
static void interp_linear(const ora_tree *t, double *dst_f, int64_t dblk, const int *dfirst,
                          const double *src_f, int64_t sblk, const int *sfirst, const int *slen,
                          int add) {
    const int D = t->D;
    const double w0 = (D == 2) ? 0.5 : 0.25;
    const double w1 = 0.25;
    int is[3] = {0, 0, 0};
    int sl[3] = {1, 1, 1};
    for (int d = 0; d < D; d++) sl[d] = slen[d];
    for (is[2] = 0; is[2] < sl[2]; is[2]++)
    for (is[1] = 0; is[1] < sl[1]; is[1]++)
    for (is[0] = 0; is[0] < sl[0]; is[0]++) {
        int Js[3], Jd0[3];
        for (int d = 0; d < D; d++) {
            Js[d] = sfirst[d] + is[d];
            Jd0[d] = dfirst[d] + 2 * is[d];
        }
        int nc = 1 << D;
        for (int c = 0; c < nc; c++) {
            int ii[3], Jd[3];
            for (int d = 0; d < D; d++) { ii[d] = (c >> d) & 1; Jd[d] = Jd0[d] + ii[d]; }
            double f = w0 * src_f[OFF(t, Js, sblk)];
            for (int j = 0; j < D; j++) {
                int K[3];
                for (int d = 0; d < D; d++) K[d] = Js[d];
                K[j] += 2 * ii[j] - 1;
                f += w1 * src_f[OFF(t, K, sblk)];
            }
            int64_t o = OFF(t, Jd, dblk);
            dst_f[o] = add ? (f + dst_f[o]) : f;
        }
    }
}

void ora_fill_ghost_interp(const ora_tree *t, double *u, int level) {
    const int D = t->D, M = t->M, G = t->G;
    const int W = 2 + 2 * D;
    const int mh = M / 2, gh = G / 2;
    int64_t lo = t->ref_off[level - 1], hi = t->ref_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        const int64_t *e = t->ref + k * W;
        int64_t fine = e[0], coarse = e[1];
        const int64_t *face = e + 2, *sb = e + 2 + D;
        int dfirst[3] = {1, 1, 1}, sfirst[3] = {1, 1, 1}, slen[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) {
            /* ghostindices(u, face) */
            dfirst[d] = (face[d] == 0) ? G + 1 : (face[d] == -1 ? 1 : M + G + 1);
            /* subblockbnd(u, sb, -face) */
            int sbf = G + 1 + (int)(sb[d] - 1) * mh;
            int64_t fd = -face[d];
            if (fd == -1) { sfirst[d] = sbf; slen[d] = gh; }
            else if (fd == 1) { sfirst[d] = sbf + mh - gh; slen[d] = gh; }
            else { sfirst[d] = sbf; slen[d] = mh; }
        }
        interp_linear(t, u, fine, dfirst, u, coarse, sfirst, slen, 0);
    }
}

void ora_fill_ghost(const ora_tree *t, double *u, int level, const int8_t *bc) {
    ora_fill_ghost_copy(t, u, level);
    ora_fill_ghost_bnd(t, u, level, bc);
    ora_fill_ghost_interp(t, u, level);
}

/* ---------------------------------------------------------------------- */
/* Smoother, residuals, transfers: src/multigrid.jl, src/restrict.jl,        */
/* src/interp.jl, src/blockfield.jl:557-576                                  */
/* ---------------------------------------------------------------------- */

/* gauss_seidel!(..., ::RedBlack{P}) (src/multigrid.jl:153-180) over a level
 * (gauss_seidel_level!, src/multigrid.jl:238-243). */
void ora_gs_level(const ora_tree *t, double *u, const double *b, double omega, double s,
                  int level, const ora_op *op, int P) {
    const int D = t->D, M = t->M, G = t->G;
    const double c = centre_coef(t, op, s);   /* diagelm */
    const double nwc = -(omega / c);
    int64_t lo = t->pairs_off[level - 1], hi = t->pairs_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        int i[3] = {1, 1, 1};
        int n2 = (D >= 2) ? M : 1, n3 = (D >= 3) ? M : 1;
        for (int a3 = 0; a3 < n3; a3++)
        for (int a2 = 0; a2 < n2; a2++)
        for (i[0] = G + 1; i[0] <= G + M; i[0] += 2) {
            i[1] = G + 1 + a2;
            i[2] = G + 1 + a3;
            int p = P;
            for (int d = 1; d < D; d++) p ^= (i[d] & 1);
            int I[3] = {i[0] + p, i[1], i[2]};
            int64_t J[3];
            global_index(t, blk, I, J);
            double Lu = applystencil(t, op, u, blk, I, J, s);
            int64_t o = OFF(t, I, blk);
            double v = fma(s, b[o], Lu);
            u[o] = fma(nwc, v, u[o]);
        }
    }
}

/* gauss_seidel!(..., ::FourColors{P1,P2}) (src/multigrid.jl:206-232) over a
 * level: I1 runs over every other cell in the first two dimensions (from the
 * first valid one), every cell in the third; in 3D the two parities shift with
 * I1[3] (1-based with ghosts), so the colours depend on G as in the reference.
 * u -= omega (Lu + s b) / c, with no fused multiply-add. */
void ora_gs4_level(const ora_tree *t, double *u, const double *b, double omega, double s,
                   int level, const ora_op *op, int P1, int P2) {
    const int D = t->D, M = t->M, G = t->G;
    const double c0 = diag4(D);
    int64_t lo = t->pairs_off[level - 1], hi = t->pairs_off[level];
    for (int64_t k = lo; k < hi; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        int n3 = (D >= 3) ? M : 1;
        for (int a3 = 0; a3 < n3; a3++)
        for (int i1 = G + 1; i1 <= G + M; i1 += 2)
        for (int i0 = G + 1; i0 <= G + M; i0 += 2) {
            int z = G + 1 + a3;
            int p1 = P1, p2 = P2;
            if (D == 3) {
                p1 = (P1 + z) % 2;
                p2 = (P2 + z) % 2;
            }
            int I[3] = {i0 + p1, i1 + p2, z};
            int64_t o = OFF(t, I, blk);
            int64_t J[3];
            global_index(t, blk, I, J);
            double Lu = applystencil(t, op, u, blk, I, J, s);
            double c = op->geom ? diag4cyl(J[op->geom - 1]) : c0;
            u[o] -= omega * (Lu + s * b[o]) / c;
        }
    }
}

/* gauss_seidel_iter! (src/multigrid.jl:250-295): order 2 red-black, order 4
 * the four colours, each colour followed by the copy and boundary ghost fills */
void ora_gs_iter(const ora_tree *t, double *u, const double *b, double omega, double s,
                 int n, int level, const ora_op *op) {
    if (op->order == 4) {
        for (int it = 0; it < n; it++)
            for (int c = 0; c < 4; c++) {
                ora_gs4_level(t, u, b, omega, s, level, op, c >> 1, c & 1);
                ora_fill_ghost_copy(t, u, level);
                ora_fill_ghost_bnd(t, u, level, op->bc);
            }
        return;
    }
    for (int it = 0; it < n; it++) {
        ora_gs_level(t, u, b, omega, s, level, op, 0);
        ora_fill_ghost_copy(t, u, level);
        ora_fill_ghost_bnd(t, u, level, op->bc);
        ora_gs_level(t, u, b, omega, s, level, op, 1);
        ora_fill_ghost_copy(t, u, level);
        ora_fill_ghost_bnd(t, u, level, op->bc);
    }
}

/* residual! / residual_level! (src/multigrid.jl:309-323, 349-354) */
void ora_residual_level(const ora_tree *t, double *r, const double *u, const double *b,
                        double s, int level, const ora_op *op) {
    const int D = t->D, M = t->M, G = t->G;
    int64_t lo = t->pairs_off[level - 1], hi = t->pairs_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        int I[3] = {1, 1, 1};
        int n2 = (D >= 2) ? M : 1, n3 = (D >= 3) ? M : 1;
        for (int a3 = 0; a3 < n3; a3++)
        for (int a2 = 0; a2 < n2; a2++)
        for (int a1 = 0; a1 < M; a1++) {
            I[0] = G + 1 + a1; I[1] = G + 1 + a2; I[2] = G + 1 + a3;
            int64_t J[3];
            global_index(t, blk, I, J);
            double Lu = applystencil(t, op, u, blk, I, J, s);
            int64_t o = OFF(t, I, blk);
            r[o] = -(b[o] + Lu / s);
        }
    }
}

/* rhs! / rhs_level! (src/multigrid.jl:363-388) over the blocks of `level`.
 * order 2 (Laplacian or Helmholtz): b1 = b (src/laplacian.jl:88-92, 121-125,
 * 296-300; src/helmholtz.jl:27-32). order 4: the compact R stencil
 * (src/laplacian.jl:142-149 2D Cartesian, 167-175 / 187-195 cylindrical,
 * 322-337 3D): the @stencil literal's nonzero terms in column-major order --
 * the order-2 stencil's order with centre weight 8 (2D) or 6 (3D) and the
 * radial weights on the radial neighbours -- then divided by 12. b's ghosts
 * of this level must be filled. */
void ora_rhs_level(const ora_tree *t, double *b1, const double *b, int level, const ora_op *op, int order) {
    const int D = t->D, M = t->M, G = t->G;
    int64_t lo = t->pairs_off[level - 1], hi = t->pairs_off[level];
    for (int64_t k = lo; k < hi; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        int I[3] = {1, 1, 1};
        int n2 = (D >= 2) ? M : 1, n3 = (D >= 3) ? M : 1;
        for (int a3 = 0; a3 < n3; a3++)
        for (int a2 = 0; a2 < n2; a2++)
        for (int a1 = 0; a1 < M; a1++) {
            I[0] = G + 1 + a1; I[1] = G + 1 + a2; I[2] = G + 1 + a3;
            int64_t o = OFF(t, I, blk);
            if (order == 4) {
                int64_t J[3];
                global_index(t, blk, I, J);
                b1[o] = stencil_sum(t, op, b, blk, I, J, (double)(12 - 2 * D)) / 12.0;
            } else {
                b1[o] = b[o];
            }
        }
    }
}

/* restrict! / restrict_level! (src/restrict.jl:5-37) over conn.child[level] */
void ora_restrict_level(const ora_tree *t, double *u, int level) {
    const int D = t->D, M = t->M, G = t->G, W = 2 + D;
    const int mh = M / 2;
    const double k = (double)(1 << D);
    int64_t lo = t->chd_off[level - 1], hi = t->chd_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t e = lo; e < hi; e++) {
        const int64_t *c = t->chd + e * W;
        int64_t fine = c[0], coarse = c[1];
        const int64_t *sb = c + 2;
        int id[3] = {0, 0, 0};
        int n2 = (D >= 2) ? mh : 1, n3 = (D >= 3) ? mh : 1;
        for (id[2] = 0; id[2] < n3; id[2]++)
        for (id[1] = 0; id[1] < n2; id[1]++)
        for (id[0] = 0; id[0] < mh; id[0]++) {
            int Js0[3] = {1, 1, 1}, Jd[3] = {1, 1, 1};
            for (int d = 0; d < D; d++) {
                Js0[d] = G + 1 + 2 * id[d];
                Jd[d] = G + 1 + (int)(sb[d] - 1) * mh + id[d];
            }
            double f = 0.0;
            int nc = 1 << D;
            for (int q = 0; q < nc; q++) {   /* i_1 innermost (@nloops order) */
                int K[3] = {1, 1, 1};
                for (int d = 0; d < D; d++) K[d] = Js0[d] + ((q >> d) & 1);
                f += u[OFF(t, K, fine)];
            }
            u[OFF(t, Jd, coarse)] = f / k;
        }
    }
}

/* residual_postrestrict_level! -> residual_subblock! (src/multigrid.jl:112-119,
 * 331-344), called as (b, u, r, 4 s_l, tree[l-1], conn.child[l]): writes the
 * FAS right-hand side into b at every coarse cell covered by a level-l block. */
void ora_postrestrict_level(const ora_tree *t, double *bout, const double *u, const double *r,
                            double s, int level, const ora_op *op) {
    const int D = t->D, M = t->M, G = t->G, W = 2 + D;
    const int mh = M / 2;
    int64_t lo = t->chd_off[level - 1], hi = t->chd_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t e = lo; e < hi; e++) {
        const int64_t *c = t->chd + e * W;
        int64_t coarse = c[1];
        const int64_t *sb = c + 2;
        int id[3] = {0, 0, 0};
        int n2 = (D >= 2) ? mh : 1, n3 = (D >= 3) ? mh : 1;
        for (id[2] = 0; id[2] < n3; id[2]++)
        for (id[1] = 0; id[1] < n2; id[1]++)
        for (id[0] = 0; id[0] < mh; id[0]++) {
            int I[3] = {1, 1, 1};
            for (int d = 0; d < D; d++) I[d] = G + 1 + (int)(sb[d] - 1) * mh + id[d];
            int64_t J[3];
            global_index(t, coarse, I, J);
            double Lu = applystencil(t, op, u, coarse, I, J, s);
            int64_t o = OFF(t, I, coarse);
            bout[o] = -(r[o] + Lu / s);
        }
    }
}

/* copyto!(dest, src, layer) (src/blockfield.jl:557-563): whole blocks, ghosts included */
void ora_copy_level(const ora_tree *t, double *dst, const double *src, int level) {
    const int D = t->D;
    int64_t lo = t->pairs_off[level - 1], hi = t->pairs_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        memcpy(dst + (blk - 1) * t->SD, src + (blk - 1) * t->SD, (size_t)t->SD * sizeof(double));
    }
}

/* diffto!(dest, a1, a2, layer) (src/blockfield.jl:569-576): dest = a1 - a2 */
void ora_diff_level(const ora_tree *t, double *dst, const double *a1, const double *a2, int level) {
    const int D = t->D;
    int64_t lo = t->pairs_off[level - 1], hi = t->pairs_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t k = lo; k < hi; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        int64_t o = (blk - 1) * t->SD;
        for (int64_t q = 0; q < t->SD; q++) dst[o + q] = a1[o + q] - a2[o + q];
    }
}

/* interp_add_level!(u, u1, conn.child[level]) (src/interp.jl:185-193) */
void ora_interp_add_level(const ora_tree *t, double *u, const double *u1, int level) {
    const int D = t->D, M = t->M, G = t->G, W = 2 + D;
    const int mh = M / 2;
    int64_t lo = t->chd_off[level - 1], hi = t->chd_off[level];
#pragma omp parallel for schedule(static)
    for (int64_t e = lo; e < hi; e++) {
        const int64_t *c = t->chd + e * W;
        int64_t fine = c[0], coarse = c[1];
        const int64_t *sb = c + 2;
        int dfirst[3] = {1, 1, 1}, sfirst[3] = {1, 1, 1}, slen[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) {
            dfirst[d] = G + 1;
            sfirst[d] = G + 1 + (int)(sb[d] - 1) * mh;
            slen[d] = mh;
        }
        interp_linear(t, u, fine, dfirst, u1, coarse, sfirst, slen, 1);
    }
}

/* ---------------------------------------------------------------------- */
/* Exact level-1 solve (documented extension; no reference counterpart).     */
/* Level 1 is the complete root grid with homogeneous mirror BCs, so its    */
/* operator is a Kronecker sum of 1D tridiagonal operators. Each 1D operator */
/* T = W^-1 S (W = diag(2i-1) on the radial axis, identity otherwise, S      */
/* symmetric) is diagonalised through A = W^-1/2 S W^-1/2 = Y L Y^T, so     */
/* T = P L Q with P = W^-1/2 Y and Q = Y^T W^1/2.                            */
/*                                                                          */
/* The eigensolver is the implicit QL method with Wilkinson shifts on the   */
/* tridiagonal A (the textbook tql2 iteration: deflate at the first         */
/* negligible off-diagonal, shift from the trailing 2x2 block, chase the    */
/* bulge with Givens rotations from the bottom up, accumulating them into  */
/* the eigenvector matrix). It is the method the HIP library uses, so both */
/* sides hold the same matrices and exact-coarse V-cycles agree bit for    */
/* bit. It is pinned by the cyclic Jacobi method below (an independent     */
/* algorithm, kept for that purpose), by the dense eigenvalues and by      */
/* scipy's spsolve of the assembled level-1 operator (tests/test_oracle.py).*/
/* ---------------------------------------------------------------------- */

/* d[n]: diagonal (eigenvalues on return); e[n]: e[i] couples i and i+1
 * (destroyed); z: n x n row-major, eigenvectors in its columns on return.
 * Returns 0, or -1 if an eigenvalue needs more than 200 iterations. */
static int ql_tridiag(int n, double *d, double *e, double *z) {
    const double eps = 2.220446049250313e-16;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) z[i * n + j] = (i == j) ? 1.0 : 0.0;
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iters = 0;
        for (;;) {
            /* first m >= l whose coupling to m+1 is negligible */
            int m;
            for (m = l; m < n - 1; m++)
                if (fabs(e[m]) <= eps * (fabs(d[m]) + fabs(d[m + 1]))) break;
            if (m == l) break;             /* d[l] has converged */
            if (++iters > 200) return -1;
            /* Wilkinson shift from the 2x2 block at l */
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double sn = 1.0, cs = 1.0, p = 0.0;
            int i;
            for (i = m - 1; i >= l; i--) {
                double f = sn * e[i], bb = cs * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {           /* exact split: restart at l */
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    break;
                }
                sn = f / r;
                cs = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * sn + 2.0 * cs * bb;
                p = sn * r;
                d[i + 1] = g + p;
                g = cs * r - bb;
                for (int k = 0; k < n; k++) {  /* rotate columns i, i+1 */
                    double zi = z[k * n + i], zi1 = z[k * n + i + 1];
                    z[k * n + i + 1] = sn * zi + cs * zi1;
                    z[k * n + i] = cs * zi - sn * zi1;
                }
            }
            if (i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

/* cyclic Jacobi: A (n x n, row-major) symmetric, destroyed; eigenvalues on its
 * diagonal; V receives the eigenvectors as columns. Only the cross-check
 * ora_axis_decomp_jacobi uses it. */
static void jacobi_eig(int n, double *A, double *V) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double a = A[i * n + j] * A[i * n + j];
                tot += a;
                if (i != j) off += a;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                double app = A[p * n + p], aqq = A[q * n + q];
                if (fabs(apq) <= 1e-300 || fabs(apq) < 1e-18 * (fabs(app) + fabs(aqq))) continue;
                double theta = (aqq - app) / (2.0 * apq);
                double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(tt * tt + 1.0), sn = tt * cs;
                for (int k = 0; k < n; k++) {          /* columns p, q */
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = cs * akp - sn * akq;
                    A[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {          /* rows p, q */
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = cs * apk - sn * aqk;
                    A[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
}

/* the 1D level-1 operator of one axis: the diagonal diag[i] of T (mirror
 * ghost folded in at the ends; the radial axis weight (2i-2)/(2i-1) is 0 at
 * i = 1), the weights w[i] of W, and the symmetric couplings
 * off[i] = S[i][i+1] / sqrt(w[i] w[i+1]) of A */
static void axis_operator(int n, int radial, double alow, double ahigh, double *diag, double *w, double *off) {
    for (int i = 0; i < n; i++) {
        int ii = i + 1;
        w[i] = radial ? (double)(2 * ii - 1) : 1.0;
        double dg = -2.0;
        if (!radial) {
            if (i == 0) dg += alow;
            if (i == n - 1) dg += ahigh;
        } else if (i == n - 1) {
            dg += ((double)(2 * ii) / (double)(2 * ii - 1)) * ahigh;
        }
        diag[i] = dg;
        off[i] = 0.0;
        if (i + 1 < n) {
            double sij = radial ? (double)(2 * ii) : 1.0;
            off[i] = sij / sqrt(w[i] * (radial ? (double)(2 * ii + 1) : 1.0));
        }
    }
}

/* 1D operator of dimension d at level 1 with n cells; returns lam[n], P, Q
 * (n x n row-major) with T = P diag(lam) Q. Returns 0, or -1 if QL fails. */
int ora_axis_decomp(int n, int radial, double alow, double ahigh, double *lam, double *P, double *Q) {
    double *w = (double *)calloc((size_t)n, sizeof(double));
    double *e = (double *)calloc((size_t)n, sizeof(double));
    double *Y = (double *)calloc((size_t)n * n, sizeof(double));
    axis_operator(n, radial, alow, ahigh, lam, w, e);
    int rc = ql_tridiag(n, lam, e, Y);
    for (int i = 0; i < n; i++) {
        double sw = sqrt(w[i]);
        for (int k = 0; k < n; k++) {
            P[(size_t)i * n + k] = Y[(size_t)i * n + k] / sw;
            Q[(size_t)k * n + i] = Y[(size_t)i * n + k] * sw;
        }
    }
    free(w); free(e); free(Y);
    return rc;
}

/* The same decomposition by the cyclic Jacobi method on the dense A
 * (cross-check of ora_axis_decomp; eigenvalues in Jacobi's order). */
void ora_axis_decomp_jacobi(int n, int radial, double alow, double ahigh, double *lam, double *P, double *Q) {
    double *w = (double *)calloc((size_t)n, sizeof(double));
    double *dg = (double *)calloc((size_t)n, sizeof(double));
    double *e = (double *)calloc((size_t)n, sizeof(double));
    double *A = (double *)calloc((size_t)n * n, sizeof(double));
    double *Y = (double *)calloc((size_t)n * n, sizeof(double));
    axis_operator(n, radial, alow, ahigh, dg, w, e);
    for (int i = 0; i < n; i++) {
        A[(size_t)i * n + i] = dg[i];
        if (i + 1 < n) A[(size_t)i * n + i + 1] = A[(size_t)(i + 1) * n + i] = e[i];
    }
    jacobi_eig(n, A, Y);
    for (int k = 0; k < n; k++) lam[k] = A[(size_t)k * n + k];
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            P[(size_t)i * n + k] = Y[(size_t)i * n + k] / sqrt(w[i]);
            Q[(size_t)k * n + i] = Y[(size_t)i * n + k] * sqrt(w[i]);
        }
    free(w); free(dg); free(e); free(A); free(Y);
}

/* apply the n x n matrix Mx along axis d of a column-major grid with sizes N[] */
static void apply_axis(const double *Mx, int d, const int *N, int D, double *x, double *tmp) {
    int64_t stride = 1;
    for (int e = 0; e < d; e++) stride *= N[e];
    int64_t total = 1;
    for (int e = 0; e < D; e++) total *= N[e];
    int n = N[d];
    int64_t outer = total / (stride * n);
    for (int64_t o = 0; o < outer; o++)
        for (int64_t s0 = 0; s0 < stride; s0++) {
            int64_t base = o * stride * n + s0;
            for (int i = 0; i < n; i++) {
                double acc = 0.0;
                for (int k = 0; k < n; k++) acc += Mx[i * n + k] * x[base + (int64_t)k * stride];
                tmp[base + (int64_t)i * stride] = acc;
            }
        }
    memcpy(x, tmp, (size_t)total * sizeof(double));
}

/* u[level 1] = exact solution of L_1 u + s b = 0, then the copy and boundary
 * ghost fills that end gauss_seidel_iter!. Returns 0, or -1 if singular. */
int ora_coarse_solve(ora_tree *t, double *u, const double *b, double s, const ora_op *op) {
    const int D = t->D, M = t->M, G = t->G;
    int nb[3] = {1, 1, 1}, N[3] = {1, 1, 1};
    /* level-1 domain extent from the pairs (level 1 is complete) */
    for (int64_t k = t->pairs_off[0]; k < t->pairs_off[1]; k++)
        for (int d = 0; d < D; d++) {
            int c = (int)t->pairs[k * (D + 1) + d];
            if (c > nb[d]) nb[d] = c;
        }
    int64_t total = 1;
    for (int d = 0; d < D; d++) { N[d] = nb[d] * M; total *= N[d]; }
    double *x = (double *)calloc((size_t)total, sizeof(double));
    double *tmp = (double *)calloc((size_t)total, sizeof(double));
    double *lam[3], *P[3], *Q[3];
    int radialdim = (op->geom == 1) ? 0 : (op->geom == 2 ? 1 : -1);
    int hit = t->ck_valid && t->ck_geom == op->geom && memcmp(t->ck_bc, op->bc, 27) == 0;
    for (int d = 0; d < D; d++) {
        if (!hit) {
            int fl[3] = {0, 0, 0}, fh[3] = {0, 0, 0};
            fl[d] = -1; fh[d] = 1;
            int64_t bl[3] = {fl[0], fl[1], fl[2]}, bh[3] = {fh[0], fh[1], fh[2]};
            double alow = (double)op->bc[bc_index(D, bl)];
            double ahigh = (double)op->bc[bc_index(D, bh)];
            free(t->ck_lam[d]); free(t->ck_P[d]); free(t->ck_Q[d]);
            t->ck_lam[d] = (double *)malloc(sizeof(double) * N[d]);
            t->ck_P[d] = (double *)malloc(sizeof(double) * N[d] * N[d]);
            t->ck_Q[d] = (double *)malloc(sizeof(double) * N[d] * N[d]);
            if (ora_axis_decomp(N[d], d == radialdim, alow, ahigh, t->ck_lam[d], t->ck_P[d], t->ck_Q[d])) {
                t->ck_valid = 0;
                free(x); free(tmp);
                return -1;
            }
        }
        lam[d] = t->ck_lam[d]; P[d] = t->ck_P[d]; Q[d] = t->ck_Q[d];
    }
    t->ck_valid = 1; t->ck_geom = op->geom;
    memcpy(t->ck_bc, op->bc, 27);
    /* gather rhs = -s b */
    for (int64_t k = t->pairs_off[0]; k < t->pairs_off[1]; k++) {
        const int64_t *p = t->pairs + k * (D + 1);
        int64_t blk = p[D];
        int a[3] = {0, 0, 0};
        int n2 = (D >= 2) ? M : 1, n3 = (D >= 3) ? M : 1;
        for (a[2] = 0; a[2] < n3; a[2]++)
        for (a[1] = 0; a[1] < n2; a[1]++)
        for (a[0] = 0; a[0] < M; a[0]++) {
            int I[3] = {G + 1 + a[0], G + 1 + a[1], G + 1 + a[2]};
            int64_t gi = 0, m = 1;
            for (int d = 0; d < D; d++) { gi += (M * (p[d] - 1) + a[d]) * m; m *= N[d]; }
            x[gi] = -s * b[OFF(t, I, blk)];
        }
    }
    for (int d = 0; d < D; d++) apply_axis(Q[d], d, N, D, x, tmp);
    double shift = op->helm ? -(s * op->k2) : 0.0;
    int singular = 0;
    for (int64_t gi = 0; gi < total; gi++) {
        int64_t r = gi;
        double ev = shift;
        for (int d = 0; d < D; d++) { ev += lam[d][r % N[d]]; r /= N[d]; }
        if (fabs(ev) < 1e-13) { singular = 1; ev = 1.0; }
        x[gi] /= ev;
    }
    /* inverse transforms in the reverse axis order (the last forward axis first) */
    for (int d = D - 1; d >= 0; d--) apply_axis(P[d], d, N, D, x, tmp);
    for (int64_t k = t->pairs_off[0]; k < t->pairs_off[1]; k++) {
        const int64_t *p = t->pairs + k * (D + 1);
        int64_t blk = p[D];
        int a[3] = {0, 0, 0};
        int n2 = (D >= 2) ? M : 1, n3 = (D >= 3) ? M : 1;
        for (a[2] = 0; a[2] < n3; a[2]++)
        for (a[1] = 0; a[1] < n2; a[1]++)
        for (a[0] = 0; a[0] < M; a[0]++) {
            int I[3] = {G + 1 + a[0], G + 1 + a[1], G + 1 + a[2]};
            int64_t gi = 0, m = 1;
            for (int d = 0; d < D; d++) { gi += (M * (p[d] - 1) + a[d]) * m; m *= N[d]; }
            u[OFF(t, I, blk)] = x[gi];
        }
    }
    free(x); free(tmp);
    ora_fill_ghost_copy(t, u, 1);
    ora_fill_ghost_bnd(t, u, 1, op->bc);
    return singular ? -1 : 0;
}

/* ---------------------------------------------------------------------- */
/* V-cycle, FMG, residual norm: src/multigrid.jl:17-106, 391-417            */
/* ---------------------------------------------------------------------- */

static inline double level_scale(double s, int l) {
    int64_t f = (int64_t)1 << (l - 1);
    return s / (double)(f * f);
}

static void top_solve(ora_tree *t, double *u, const double *b, double omega, double s,
                      int ntop, const ora_op *op, int coarse_exact) {
    if (coarse_exact) ora_coarse_solve(t, u, b, s, op);
    else ora_gs_iter(t, u, b, omega, s, ntop, 1, op);
}

/* vcycle! (src/multigrid.jl:60-106). coarse_exact = 0 is the reference's
 * ntop Gauss-Seidel sweeps at level 1. */
void ora_vcycle(ora_tree *t, double *u, double *b, double *r, double *u1, double s,
                const ora_op *op, int lmax, int ndown, int ntop, int nup, double omega,
                int coarse_exact) {
    if (lmax == 0) lmax = t->nlevels;
    for (int l = lmax; l >= 2; l--) {
        double sl = level_scale(s, l);
        ora_gs_iter(t, u, b, omega, sl, ndown, l, op);
        ora_restrict_level(t, u, l);
        ora_fill_ghost(t, u, l - 1, op->bc);
        ora_copy_level(t, u1, u, l - 1);
        ora_residual_level(t, r, u, b, sl, l, op);
        ora_fill_ghost(t, r, l, op->bc);
        ora_restrict_level(t, r, l);
        ora_fill_ghost(t, r, l - 1, op->bc);
        ora_postrestrict_level(t, b, u, r, 4 * sl, l, op);
    }
    top_solve(t, u, b, omega, s, ntop, op, coarse_exact);
    for (int l = 2; l <= lmax; l++) {
        double sl = level_scale(s, l);
        ora_diff_level(t, u1, u, u1, l - 1);
        ora_fill_ghost(t, u1, l - 1, op->bc);
        ora_interp_add_level(t, u, u1, l);
        ora_fill_ghost(t, u, l, op->bc);
        ora_gs_iter(t, u, b, omega, sl, nup, l, op);
        ora_fill_ghost(t, u, l, op->bc);
    }
}

static int finest_level(const ora_tree *t) {
    int last = 0;
    for (int l = 0; l < t->nlevels; l++)
        if (t->pairs_off[l + 1] > t->pairs_off[l]) last = l + 1;
    return last;
}

/* fmg! (src/multigrid.jl:17-51) */
void ora_fmg(ora_tree *t, double *u, double *b, double *r, double *u1, double s,
             const ora_op *op, int ndown, int ntop, int nup, double omega, int coarse_exact) {
    int lmax = finest_level(t);
    for (int l = lmax; l >= 2; l--) {
        double sl = level_scale(s, l);
        ora_restrict_level(t, u, l);
        ora_fill_ghost(t, u, l - 1, op->bc);
        ora_residual_level(t, r, u, b, sl, l, op);
        ora_fill_ghost(t, r, l, op->bc);
        ora_restrict_level(t, r, l);
        ora_fill_ghost(t, r, l - 1, op->bc);
        ora_postrestrict_level(t, b, u, r, 4 * sl, l, op);
    }
    for (int l = 1; l <= lmax; l++) {
        ora_copy_level(t, u1, u, l);
        if (l > 1) {
            ora_diff_level(t, u1, u, u1, l - 1);
            ora_fill_ghost(t, u1, l - 1, op->bc);
            ora_interp_add_level(t, u, u1, l);
        }
        ora_fill_ghost(t, u, l, op->bc);
        ora_vcycle(t, u, b, r, u1, s, op, l, ndown, ntop, nup, omega, coarse_exact);
    }
}

/* resnorm! (src/multigrid.jl:391-417). Overwrites r. */
double ora_resnorm(const ora_tree *t, double *r, const double *b, const double *u, double s,
                   const ora_op *op) {
    const int D = t->D, M = t->M, G = t->G;
    int lmax = t->nlevels;
    for (int l = lmax; l >= 1; l--) ora_residual_level(t, r, u, b, level_scale(s, l), l, op);
    int64_t tot = t->nblocks * t->SD;
    for (int64_t q = 0; q < tot; q++) r[q] = r[q] * r[q];
    for (int l = lmax - 1; l >= 1; l--) ora_restrict_level(t, r, l + 1);
    double w = 0.0;
    for (int64_t k = t->pairs_off[0]; k < t->pairs_off[1]; k++) {
        int64_t blk = t->pairs[k * (D + 1) + D];
        double part = 0.0;
        int first = 1;
        int I[3] = {1, 1, 1};
        int n2 = (D >= 2) ? M : 1, n3 = (D >= 3) ? M : 1;
        for (int a3 = 0; a3 < n3; a3++)
        for (int a2 = 0; a2 < n2; a2++)
        for (int a1 = 0; a1 < M; a1++) {
            I[0] = G + 1 + a1; I[1] = G + 1 + a2; I[2] = G + 1 + a3;
            double v = r[OFF(t, I, blk)];
            if (first) { part = v; first = 0; } else part += v;
        }
        w += part;
    }
    return sqrt(w);
}

/* Caller-side pre-solve fill: fill_ghost!(u, l) for l = 1..L (src/streamer.jl:323-325) */
void ora_fill_ghost_all(const ora_tree *t, double *u, const int8_t *bc) {
    for (int l = 1; l <= t->nlevels; l++) ora_fill_ghost(t, u, l, bc);
}

/* ---------------------------------------------------------------------- */
/* electric_field! (src/multigrid.jl:421-449) with a constant external     */
/* field e0, the method of src/multigrid.jl:451                             */
/* ---------------------------------------------------------------------- */

/* For every block of every level (the @bkernel loop), with h_l = h / 2^(l-1):
 *   f[I, d, blk] = (u[I - e_d, blk] - u[I, blk]) / h_l + e0[d]
 * for I in validindices(f, d) (src/blockfield.jl:403-405: G+1..G+M+1 along d,
 * G+1..G+M elsewhere), then for I in validindices(fabs)
 *   fabs[I, blk] = sqrt(sum_d (f[I, d, blk] + f[I + e_d, d, blk])^2 / 4).
 * f is a VectorBlockField{D,M,G} with contiguous storage
 * (src/blockfield.jl:88-95): (M+2G+1)^D x D x nblocks, column-major. u must have
 * its ghosts filled as at src/streamer.jl:340-349. */
void ora_electric_field(const ora_tree *t, double *f, double *fabs, const double *u, double h,
                        const double *e0) {
    const int D = t->D, M = t->M, G = t->G;
    const int64_t Sv = M + 2 * G + 1;
    int64_t SvD = 1;
    for (int d = 0; d < D; d++) SvD *= Sv;
    for (int l = 1; l <= t->nlevels; l++) {
        const double hl = h / (double)(1 << (l - 1));
        for (int64_t k = t->pairs_off[l - 1]; k < t->pairs_off[l]; k++) {
            const int64_t blk = t->pairs[k * (D + 1) + D];
            double *fb = f + (blk - 1) * D * SvD;
            for (int d = 0; d < D; d++) {
                int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
                for (int e = 0; e < D; e++) {
                    lo[e] = G + 1;
                    hi[e] = (e == d) ? G + M + 1 : G + M;
                }
                int I[3];
                for (I[2] = lo[2]; I[2] <= hi[2]; I[2]++)
                    for (I[1] = lo[1]; I[1] <= hi[1]; I[1]++)
                        for (I[0] = lo[0]; I[0] <= hi[0]; I[0]++) {
                            int I1[3] = {I[0], I[1], I[2]};
                            I1[d] -= 1;
                            int64_t fo = 0, st = 1;
                            for (int e = 0; e < D; e++) {
                                fo += (int64_t)(I[e] - 1) * st;
                                st *= Sv;
                            }
                            fb[fo + d * SvD] = (u[OFF(t, I1, blk)] - u[OFF(t, I, blk)]) / hl + e0[d];
                        }
            }
            int I[3] = {1, 1, 1};
            int lo = G + 1, hiv = G + M;
            for (I[2] = (D == 3 ? lo : 1); I[2] <= (D == 3 ? hiv : 1); I[2]++)
                for (I[1] = lo; I[1] <= hiv; I[1]++)
                    for (I[0] = lo; I[0] <= hiv; I[0]++) {
                        double f2 = 0.0;
                        for (int d = 0; d < D; d++) {
                            int64_t fo = 0, fo1, st = 1;
                            for (int e = 0; e < D; e++) {
                                fo += (int64_t)(I[e] - 1) * st;
                                st *= Sv;
                            }
                            st = 1;
                            for (int e = 0; e < d; e++) st *= Sv;
                            fo1 = fo + st;
                            const double a = fb[fo + d * SvD] + fb[fo1 + d * SvD];
                            f2 += a * a / 4.0;
                        }
                        fabs[OFF(t, I, blk)] = sqrt(f2);
                    }
        }
    }
}

/* ---- free boundary conditions (src/freebc.jl; cylindrical 2D) ----
 * dim is the geometry's radial dimension (1-based, CylindricalGeometry{dim});
 * rootsize[D] the level-1 block domain; n = rootsize[other] 2^(level-1) M. */

/* block of level l at coordinate (kdim, iother), 0 if absent */
static int64_t block_at(const ora_tree *t, int l, int dim, int64_t kdim, int64_t iother) {
    const int D = t->D, od = (dim % D) + 1;
    for (int64_t k = t->pairs_off[l - 1]; k < t->pairs_off[l]; k++) {
        const int64_t *p = t->pairs + k * (D + 1);
        if (p[dim - 1] == kdim && p[od - 1] == iother) return p[D];
    }
    return 0;
}

/* getouter!(u1, u, tree, dim, level, h) (src/freebc.jl:112-142) */
void ora_freebc_getouter(const ora_tree *t, double *u1, const double *u, int dim, int level, double h,
                         const int64_t *rootsize) {
    const int D = t->D, M = t->M, G = t->G, od = (dim % D) + 1;
    for (int l = 1; l <= level && l <= t->nlevels; l++) {
        const int64_t kd = rootsize[dim - 1] << (l - 1), no = rootsize[od - 1] << (l - 1);
        const double hl = h / (double)(1 << (l - 1));
        const int64_t p = (int64_t)1 << (level - l);
        for (int64_t iblk = 1; iblk <= no; iblk++) {
            const int64_t blk = block_at(t, l, dim, kd, iblk);
            if (!blk) continue;
            const int64_t i0 = (iblk - 1) * M;
            for (int i = 1; i <= M; i++) {
                int I[3] = {1, 1, 1};
                I[dim - 1] = G + M;
                I[od - 1] = G + i;
                const double v = u[OFF(t, I, blk)] / hl;
                for (int64_t j = p * (i0 + i - 1); j < p * (i0 + i); j++) u1[j] = v;
            }
        }
    }
}

/* FFTW's RODFT10 (DST-II), y_k = 2 sum_j x_j sin(pi (j + 1/2)(k + 1)/n), and
 * RODFT01 (DST-III), y_k = (-1)^k x_{n-1} + 2 sum_{j<n-1} x_j sin(pi (j + 1)(k + 1/2)/n),
 * as direct sums in j order (the FFTW plans of src/freebc.jl:45-46) */
void ora_dst_matrix(double *S, int64_t n, int type) {
    for (int64_t k = 0; k < n; k++)
        for (int64_t j = 0; j < n; j++) {
            double v;
            if (type == 2) v = 2.0 * sin(M_PI * ((double)j + 0.5) * (double)(k + 1) / (double)n);
            else v = (j == n - 1) ? ((k & 1) ? -1.0 : 1.0)
                                   : 2.0 * sin(M_PI * (double)(j + 1) * ((double)k + 0.5) / (double)n);
            S[k * n + j] = v;
        }
}

static void matvec(const double *S, const double *x, double *y, int64_t n) {
    for (int64_t k = 0; k < n; k++) {
        double acc = 0.0;
        for (int64_t j = 0; j < n; j++) acc += S[k * n + j] * x[j];
        y[k] = acc;
    }
}

/* sum of b[a .. a+len) (0-based) in order */
static double bsum(const double *b, int64_t a, int64_t len) {
    double s = 0.0;
    for (int64_t j = a; j < a + len; j++) s += b[j];
    return s;
}

/* factor(::CylindricalGeometry{dim}, c, delta) with delta = e_dim (src/geometry.jl:27-29) */
static double cyl_factor(int64_t cdim) { return fabs(1.0 + 1.0 / (double)(2 * cdim - 1)); }

/* setouter!(q, b, tree, conn, dim, level, s, geom) (src/freebc.jl:149-215) */
void ora_freebc_setouter(const ora_tree *t, double *q, const double *b, int dim, int level, double s,
                         const int64_t *rootsize) {
    const int D = t->D, M = t->M, G = t->G, od = (dim % D) + 1;
    const int W = 2 + 2 * D;
    for (int l = 1; l <= level && l <= t->nlevels; l++) {
        const double sl = s / (double)(((int64_t)1 << (l - 1)) * ((int64_t)1 << (l - 1)));
        const int64_t p = (int64_t)1 << (level - l);
        const int64_t kd = rootsize[dim - 1] << (l - 1);
        for (int64_t k = t->bnd_off[l - 1]; k < t->bnd_off[l]; k++) {
            const int64_t *e = t->bnd + k * W;
            const int64_t *face = e + 1;
            int is0 = 1;
            for (int d = 0; d < D; d++) is0 &= (face[d] == (d == dim - 1 ? 1 : 0));
            if (!is0) continue;
            const int64_t iblk = e[0];
            const int64_t *B = t->coord_of + (iblk - 1) * D;
            const int64_t i0 = (B[od - 1] - 1) * M;
            for (int i = 1; i <= M; i++) {
                int I[3] = {1, 1, 1};
                I[dim - 1] = G + M;
                I[od - 1] = G + i;
                const int64_t Jd = M * (B[dim - 1] - 1) + M;
                const double bmean = bsum(b, p * (i0 + i - 1), p) / (double)p;
                q[OFF(t, I, iblk)] += 2.0 * cyl_factor(Jd) * bmean / sl;
            }
        }
        for (int64_t k = t->ref_off[l - 1]; k < t->ref_off[l]; k++) {
            const int64_t *e = t->ref + k * W;
            const int64_t iblk = e[0];
            const int64_t *face = e + 2;
            const int64_t *B = t->coord_of + (iblk - 1) * D;
            if (B[dim - 1] != kd) continue;
            const int64_t dir = face[od - 1];
            const int i = (dir == -1) ? G + 1 : M + G;
            int I[3] = {1, 1, 1};
            I[dim - 1] = G + M;
            I[od - 1] = i;
            const int64_t j = (B[od - 1] - 1) * M + i - G;
            const int64_t Jd = M * (B[dim - 1] - 1) + M;
            const int64_t J1d = (Jd - 1) / 2 + 1;
            double bmean;
            if (dir == 1) bmean = bsum(b, p * j, 2 * p) / (double)(2 * p);
            else bmean = bsum(b, p * (j - 3) - 1, 2 * p) / (double)(2 * p);
            q[OFF(t, I, iblk)] += 2.0 * 0.25 * cyl_factor(J1d) * bmean / sl;
        }
    }
}

/* freebc!(q, u, h, s, tree, conn, conf) (src/freebc.jl:87-110): f[n] is the
 * kernel of the FreeBC constructor (src/freebc.jl:32-44); b_out[n] receives b */
void ora_freebc(const ora_tree *t, double *q, const double *u, double h, double s, int dim, int level,
                const int64_t *rootsize, const double *f, int64_t n, double *b_out) {
    double *S = (double *)malloc((size_t)(n * n) * sizeof(double));
    double *u1 = (double *)calloc((size_t)n, sizeof(double));
    double *am = (double *)malloc((size_t)n * sizeof(double));
    ora_freebc_getouter(t, u1, u, dim, level, h, rootsize);
    ora_dst_matrix(S, n, 2);
    matvec(S, u1, am, n);
    /* am[end] *= 3 acts on the previous contents of am, which the next
     * line overwrites (src/freebc.jl:102-103) */
    for (int64_t k = 0; k < n; k++) am[k] = f[k] * am[k];
    ora_dst_matrix(S, n, 3);
    matvec(S, am, u1, n);
    for (int64_t k = 0; k < n; k++) b_out[k] = u1[k] * h / (double)(1 << (level - 1)) / 2.0;
    ora_freebc_setouter(t, q, b_out, dim, level, s, rootsize);
    free(S); free(u1); free(am);
}

/* fill_ghost_free!(u, conf, tree, l) (src/freebc.jl:218-238) */
void ora_fill_ghost_free(const ora_tree *t, double *u, const double *b, int dim, int level, int l,
                         const int64_t *rootsize) {
    const int D = t->D, M = t->M, G = t->G, od = (dim % D) + 1;
    const int64_t kd = rootsize[dim - 1] << (l - 1), no = rootsize[od - 1] << (l - 1);
    const int64_t p = (int64_t)1 << (level - l);
    for (int64_t iblk = 1; iblk <= no; iblk++) {
        const int64_t blk = block_at(t, l, dim, kd, iblk);
        if (!blk) continue;
        const int64_t i0 = (iblk - 1) * M;
        for (int i = 1; i <= M; i++) {
            int I[3] = {1, 1, 1};
            I[dim - 1] = G + M + 1;
            I[od - 1] = G + i;
            const double bmean = bsum(b, p * (i0 + i - 1), p) / (double)p;
            u[OFF(t, I, blk)] += 2.0 * bmean;
        }
    }
}
