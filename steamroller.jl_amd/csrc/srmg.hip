// libsrmg: MI355X-native block-structured AMR geometric multigrid.
//
// Host side: context and regrid tables, host<->device layout conversion,
// and the orchestration of vcycle!/fmg!/resnorm! (src/multigrid.jl) as kernel
// sequences on one HIP stream. The C ABI is declared in include/srmg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/srmg.h"
#include "common.h"
#include "kernels.h"
#include "layout.h"
#include "unet.h"

using namespace srmg;

// first-stage work groups of the fixed-order sum of partials
constexpr int SUM_CHUNKS = 256;
// per-tree totals start here in d_sum
constexpr int TSUM = SUM_CHUNKS + 4;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

struct FieldBuf {
    int dtype = SRMG_F64;
    DevBuf data;  // nb * V values
    DevBuf alt;   // second buffer for ping-pong smoothing (made on first sweep)
    DevBuf ib;    // nslot * F values (coarse-fine ghost slabs)
    DevBuf frame; // order 4: every block's depth-1 ghost frame (made on first use)
    std::vector<uint8_t> cur;  // per level: 0 -> values in data, 1 -> in alt
    size_t word() const { return dtype == SRMG_F64 ? 8 : 4; }
    bool settled() const {
        for (auto v : cur)
            if (v) return false;
        return true;
    }
};

struct LevelInfo {
    int blk0 = 0, nblk = 0;     // device blocks
    int slot0 = 0, nslot = 0;   // coarse-fine ghost slots owned by this level's blocks
    int leaf0 = 0, nleaf = 0;   // entries in the leaf list
    int ext[3] = {1, 1, 1};     // largest level domain in blocks over the trees
    int wofs[2] = {0, 0};       // radial weight table offsets (dims 0 and 1)
    bool grouped = false;       // blocks form complete sibling groups of 2^D (k_gs_sweep_group)
    int gord0 = -1;             // the level's group processing order in d_gord, or -1
    bool all_leaves = false;    // every block is a leaf and the leaf list is the block order
};

// One tree of a forest: an independent problem whose blocks share the
// level-by-level launches with the other trees.
struct TreeInfo {
    int64_t host_off = 0;       // its host block ids are host_off + (1 .. nblocks)
    int64_t nhost = 0;
    int ext1[3] = {1, 1, 1};    // level-1 domain in blocks
    int group = 0, gslot = 0;   // coarse-solve group and position in it
    std::vector<int> blk0, nblk, leafk0, nleaf;  // per level: block range, position in the level's leaf list
};

// Trees whose level-1 grids have the same shape share one eigendecomposition
// and one batched exact coarse solve.
struct CoarseGroup {
    int N[3] = {1, 1, 1};
    int ntree = 0;
    int l1_blk0 = 0, nper = 0;  // level-1 device blocks: contiguous, nper per tree
    bool valid = false;
    int geom = -1;
    int8_t bc[27] = {0};
    DevBuf P[3], Q[3], lam[3];
    DevBuf cs, shift;           // per tree: -s (gather factor) and the Helmholtz shift -s k2
    DevBuf btab;                // per tree, the level-1 block at block position (bz nb1 + by) nb0 + bx
};

// a captured solve step: the graph and its executable
struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    uint64_t gen = 0;     // srmg_ctx::graph_gen when captured
    bool ran = false;     // the step has run eagerly once (lazily built state exists)
    ~Graph() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

struct srmg_ctx {
    int D = 2, M = 16, V = 0, F = 0;
    int nlevels = 0;
    int64_t nhost = 0;      // host block ids 1..nhost (all trees)
    int nb = 0;             // device blocks
    int nslot = 0;
    std::vector<LevelInfo> lv;
    std::vector<TreeInfo> trees;
    std::vector<CoarseGroup> groups;
    // host copies of the tables
    std::vector<int> h_nbr, h_coord, h_parent, h_sub, h_level, h_leaves, h_tree;
    std::vector<int> host2dev, dev2host;
    std::vector<int4> h_slots;
    // device tables
    DevBuf d_nbr, d_coord, d_parent, d_sub, d_slots, d_leaves, d_halo, d_isleaf, d_pix;
    DevBuf d_w64, d_w32;    // radial weights (double / float), per level and dim
    DevBuf d_partial, d_sum;
    DevBuf d_rng_leaf, d_rng_blk;  // per (tree, level): ranges of the leaf / block partials
    DevBuf d_gord;                 // per grouped 3D level: processing order of its sibling groups
    DevBuf d_grid, d_grid2; // stacked level-1 uniform grids (coarse solve)
    std::vector<std::unique_ptr<FieldBuf>> fields;
    std::unique_ptr<FieldBuf> zero[2];  // all-zero scratch fields (F64, F32), made on demand
    srmg_unet *guess_net = nullptr;     // network of the learned first iterate (not owned)
    bool group_tiles = true;            // sibling-group kernels on grouped levels
    int axis_gemm = 2;                  // level-1 transform kernels (launch_axis_gemm, launch_l1_fused)
    int group_order = 0;                // column sweeps: groups in block order (0) or slab order (1)
    int guess_fallbacks = 0;            // trees of the last solve that started from zero instead
    hipStream_t stream = nullptr;
    int maxpartials = 0;
    std::vector<int> sweep_cfg;  // per (dtype, level): tile shape of the fused sweep (see Ops::sweep)
    // the current problem of every tree (prepare()): scale s and operator
    std::vector<double> ps;
    std::vector<srmg_op> pops;
    double pomega = 1.0;
    std::vector<double> pp_host;   // 3 x ntrees x nlevels: (s_l, cc, nwc)
    DevBuf d_pp64, d_pp32;
    // captured solve steps, keyed by everything their launches bake in (solve
    // kind, fields, scales, operators, options, tile shapes, kernel switches);
    // cleared whenever a buffer they point to may move or change its role
    std::unordered_map<std::string, std::unique_ptr<Graph>> graphs;
    uint64_t graph_gen = 0;
    // this device's copy of the file-scope scratch array (srmg_grid_solve_noalloc)
    void *scratch = nullptr;
    // the order-4 operator's box neighbour table and coarse-fine slots (built
    // on first use; see box_table)
    DevBuf d_box, d_box_slots;
    bool box_ok = false;  // bumped when coarse-solve matrices are rebuilt (older graphs point to freed ones)
};

static void alloc_field(srmg_ctx *c, FieldBuf &f);

static FieldBuf &zero_field(srmg_ctx *c, int dtype) {
    auto &z = c->zero[dtype];
    if (!z || z->data.bytes != (size_t)c->nb * c->V * (dtype == SRMG_F64 ? 8 : 4)) {
        z = std::make_unique<FieldBuf>();
        z->dtype = dtype;
        alloc_field(c, *z);
    }
    return *z;
}

static FieldBuf &field(srmg_ctx *c, int id) {
    if (id < 0 || id >= (int)c->fields.size() || !c->fields[id]) fail("invalid field id " + std::to_string(id));
    return *c->fields[id];
}

static void alloc_field(srmg_ctx *c, FieldBuf &f) {
    f.frame.release();
    f.cur.assign(c->nlevels, 0);  // alt (ping-pong) holds nothing until the next sweep
    f.data.alloc((size_t)c->nb * c->V * f.word());
    f.ib.alloc((size_t)std::max(c->nslot, 1) * c->F * f.word());
    HIPCHK(hipMemsetAsync(f.data.p, 0, f.data.bytes, c->stream));
    HIPCHK(hipMemsetAsync(f.ib.p, 0, f.ib.bytes, c->stream));
}

// ---------------------------------------------------------------------------
// regrid: device tables from the reference's records
// ---------------------------------------------------------------------------

// bit i of x to bit 3i (i < 21) and to bit 2i (i < 32)
static inline uint64_t spread3(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
static inline uint64_t spread2(uint64_t x) {
    x &= 0xffffffffull;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

// Morton code of a 1-based block coordinate: bit b of (c[d] - 1) at bit b D + d
// (21 bits per dimension in 3D)
static uint64_t morton(const int64_t *c, int D) {
    if (D == 3) return spread3(c[0] - 1) | spread3(c[1] - 1) << 1 | spread3(c[2] - 1) << 2;
    return spread2((uint64_t)(c[0] - 1) & 0x1fffff) | spread2((uint64_t)(c[1] - 1) & 0x1fffff) << 1;
}

// One tree's records as the reference lays them out (src/connectivity.jl:6-42),
// concatenated over its levels with nlevels+1 offsets.
struct TreeRecords {
    int nlevels;
    int64_t nblocks;
    const int64_t *pairs, *pairs_off, *nbr, *nbr_off, *bnd, *bnd_off, *ref, *ref_off, *chd, *chd_off;
};

// The host tables of a forest, built and checked into `c`, a staging context
// that holds only D, M and V: nothing of the caller's context changes here,
// so records that fail a check leave the context with its old tree and fields.
static void build_tables(srmg_ctx *c, const std::vector<TreeRecords> &tr) {
    const int D = c->D, M = c->M, ND = 2 * D;
    const int nt = (int)tr.size();
    if (nt < 1) fail("a forest needs at least one tree");
    int nlevels = 0;
    int64_t nhost = 0;
    for (auto &t : tr) {
        if (t.nlevels < 1) fail("nlevels must be >= 1");
        nlevels = std::max(nlevels, t.nlevels);
        nhost += t.nblocks;
    }
    c->nlevels = nlevels;
    c->nhost = nhost;
    c->trees.assign(nt, TreeInfo());
    c->lv.assign(nlevels, LevelInfo());
    // level-1 extents and coarse groups (same level-1 shape)
    c->groups.clear();
    int64_t off = 0;
    for (int t = 0; t < nt; t++) {
        TreeInfo &T = c->trees[t];
        T.host_off = off;
        T.nhost = tr[t].nblocks;
        off += tr[t].nblocks;
        for (int64_t k = tr[t].pairs_off[0]; k < tr[t].pairs_off[1]; k++)
            for (int d = 0; d < D; d++) T.ext1[d] = std::max(T.ext1[d], (int)tr[t].pairs[k * (D + 1) + d]);
        int64_t n1 = 1;
        for (int d = 0; d < D; d++) n1 *= T.ext1[d];
        if (tr[t].pairs_off[1] - tr[t].pairs_off[0] != n1) fail("level 1 must be complete (populate!)");
        int g = -1;
        for (int k = 0; k < (int)c->groups.size(); k++)
            if (std::equal(T.ext1, T.ext1 + D, c->groups[k].N)) g = k;
        if (g < 0) {
            c->groups.emplace_back();
            g = (int)c->groups.size() - 1;
            for (int d = 0; d < 3; d++) c->groups[g].N[d] = T.ext1[d];  // in blocks for now
        }
        T.group = g;
        T.gslot = c->groups[g].ntree++;
    }
    // device order: by level, then coarse group, then tree, then Morton rank
    std::vector<int> torder(nt);
    for (int t = 0; t < nt; t++) torder[t] = t;
    std::stable_sort(torder.begin(), torder.end(),
                     [&](int a, int b) { return c->trees[a].group < c->trees[b].group; });
    int nb = 0;
    for (auto &t : tr) nb += (int)t.pairs_off[t.nlevels];
    c->nb = nb;
    c->host2dev.assign(nhost + 1, -1);
    c->dev2host.assign(nb, 0);
    c->h_coord.assign((size_t)nb * D, 0);
    c->h_level.assign(nb, 0);
    c->h_tree.assign(nb, 0);
    for (auto &T : c->trees) {
        T.blk0.assign(nlevels, 0);
        T.nblk.assign(nlevels, 0);
        T.leafk0.assign(nlevels, 0);
        T.nleaf.assign(nlevels, 0);
    }
    int dev = 0;
    for (int l = 0; l < nlevels; l++) {
        LevelInfo &L = c->lv[l];
        L.blk0 = dev;
        for (int t : torder) {
            TreeInfo &T = c->trees[t];
            for (int d = 0; d < D; d++) L.ext[d] = std::max(L.ext[d], T.ext1[d] << l);
            T.blk0[l] = dev;
            if (l >= tr[t].nlevels) continue;
            std::vector<std::pair<uint64_t, int64_t>> order;
            for (int64_t k = tr[t].pairs_off[l]; k < tr[t].pairs_off[l + 1]; k++)
                order.push_back({morton(tr[t].pairs + k * (D + 1), D), k});
            std::sort(order.begin(), order.end());
            for (auto &o : order) {
                const int64_t *p = tr[t].pairs + o.second * (D + 1);
                int64_t lid = p[D];
                if (lid < 1 || lid > T.nhost) fail("block id out of range");
                int64_t hid = T.host_off + lid;
                if (c->host2dev[hid] >= 0) fail("block id appears twice");
                c->host2dev[hid] = dev;
                c->dev2host[dev] = (int)hid;
                for (int d = 0; d < D; d++) {
                    int cd = (int)p[d];
                    if (cd < 1 || cd > (T.ext1[d] << l)) fail("block coordinate outside the level domain");
                    c->h_coord[(size_t)dev * D + d] = cd;
                }
                c->h_level[dev] = l;
                c->h_tree[dev] = t;
                dev++;
            }
            T.nblk[l] = dev - T.blk0[l];
        }
        L.nblk = dev - L.blk0;
    }
    // sibling groups: 2^D consecutive blocks of one tree whose coordinates are
    // 2P + (bit d of the position in the group)
    const int NBg = 1 << D;
    for (int l = 0; l < nlevels; l++) {
        LevelInfo &L = c->lv[l];
        bool ok = L.nblk > 0 && L.nblk % NBg == 0;
        for (int g = 0; ok && g < L.nblk / NBg; g++) {
            const int b0 = L.blk0 + g * NBg;
            for (int k = 0; ok && k < NBg; k++) {
                const int bk = b0 + k;
                if (c->h_tree[bk] != c->h_tree[b0]) ok = false;
                for (int d = 0; ok && d < D; d++) {
                    const int c0 = c->h_coord[(size_t)b0 * D + d] - 1, ck = c->h_coord[(size_t)bk * D + d] - 1;
                    if ((c0 & 1) != 0 || ck != c0 + ((k >> d) & 1)) ok = false;
                }
            }
        }
        L.grouped = ok;
    }
    for (auto &G : c->groups) {
        int n1 = 1;
        for (int d = 0; d < D; d++) n1 *= G.N[d];
        G.nper = n1;
    }
    for (int t : torder) {
        CoarseGroup &G = c->groups[c->trees[t].group];
        if (c->trees[t].gslot == 0) G.l1_blk0 = c->trees[t].blk0[0];
    }
    for (auto &G : c->groups)
        for (int d = 0; d < D; d++) G.N[d] *= M;  // now in cells

    auto face_code = [&](const int64_t *f) {
        int code = -1;
        for (int d = 0; d < D; d++)
            if (f[d] != 0) {
                if (code >= 0) fail("only star-stencil faces are supported (order 2)");
                code = 2 * d + (f[d] > 0 ? 1 : 0);
            }
        if (code < 0) fail("zero face");
        return code;
    };
    c->h_nbr.assign((size_t)nb * ND, INT32_MIN);
    c->h_parent.assign(nb, -1);
    c->h_sub.assign(nb, 0);
    std::vector<std::vector<std::pair<int64_t, int4>>> slots_by_level(nlevels);
    for (int t = 0; t < nt; t++) {
        const TreeRecords &R = tr[t];
        const int64_t hoff = c->trees[t].host_off;
        auto d_of = [&](int64_t lid) {
            if (lid < 1 || lid > c->trees[t].nhost || c->host2dev[hoff + lid] < 0)
                fail("record names an unknown block");
            return c->host2dev[hoff + lid];
        };
        for (int l = 0; l < R.nlevels; l++) {
            // Neighbor(from, to, f): `to`'s ghost on face -f comes from `from`
            for (int64_t k = R.nbr_off[l]; k < R.nbr_off[l + 1]; k++) {
                const int64_t *e = R.nbr + k * (2 + D);
                int from = d_of(e[0]), to = d_of(e[1]);
                c->h_nbr[(size_t)to * ND + (face_code(e + 2) ^ 1)] = from;
            }
            for (int64_t k = R.bnd_off[l]; k < R.bnd_off[l + 1]; k++) {
                const int64_t *e = R.bnd + k * (2 + 2 * D);
                c->h_nbr[(size_t)d_of(e[0]) * ND + face_code(e + 1)] = NB_BOUNDARY;
            }
            for (int64_t k = R.ref_off[l]; k < R.ref_off[l + 1]; k++) {
                const int64_t *e = R.ref + k * (2 + 2 * D);
                int fine = d_of(e[0]), coarse = d_of(e[1]);
                int f = face_code(e + 2);
                int sb = 0;
                for (int d = 0; d < D; d++) {
                    int64_t sv = e[2 + D + d];
                    if (sv != 1 && sv != 2) fail("bad subblock");
                    sb |= (int)(sv - 1) << d;
                }
                if (c->h_level[fine] != l || c->h_level[coarse] != l - 1) fail("refinement boundary level mismatch");
                slots_by_level[l].push_back({(int64_t)fine * ND + f, make_int4(fine, f, coarse, sb)});
            }
            for (int64_t k = R.chd_off[l]; k < R.chd_off[l + 1]; k++) {
                const int64_t *e = R.chd + k * (2 + D);
                int fine = d_of(e[0]), coarse = d_of(e[1]);
                int sb = 0;
                for (int d = 0; d < D; d++) sb |= (int)(e[2 + d] - 1) << d;
                c->h_parent[fine] = coarse;
                c->h_sub[fine] = sb;
            }
        }
    }
    // coarse-fine ghost slots, grouped by the level of the fine block
    c->h_slots.clear();
    for (int l = 0; l < nlevels; l++) {
        c->lv[l].slot0 = (int)c->h_slots.size();
        auto &tmp = slots_by_level[l];
        std::sort(tmp.begin(), tmp.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (auto &t : tmp) {
            int k = (int)c->h_slots.size();
            c->h_nbr[(size_t)t.second.x * ND + t.second.y] = -2 - k;
            c->h_slots.push_back(t.second);
        }
        c->lv[l].nslot = (int)c->h_slots.size() - c->lv[l].slot0;
    }
    c->nslot = (int)c->h_slots.size();
    for (size_t i = 0; i < c->h_nbr.size(); i++)
        if (c->h_nbr[i] == INT32_MIN)
            fail("block " + std::to_string(c->dev2host[i / ND]) + " face " + std::to_string(i % ND) +
                 " has no neighbour, boundary or refinement record");
    for (int l = 1; l < nlevels; l++)
        for (int b = c->lv[l].blk0; b < c->lv[l].blk0 + c->lv[l].nblk; b++)
            if (c->h_parent[b] < 0) fail("block without Child record");
    // leaves: blocks without children; per level ordered like the blocks (tree by tree)
    std::vector<char> has_child(nb, 0);
    for (int b = 0; b < nb; b++)
        if (c->h_parent[b] >= 0) has_child[c->h_parent[b]] = 1;
    c->h_leaves.clear();
    for (int l = 0; l < nlevels; l++) {
        c->lv[l].leaf0 = (int)c->h_leaves.size();
        for (int t : torder) {
            TreeInfo &T = c->trees[t];
            T.leafk0[l] = (int)c->h_leaves.size() - c->lv[l].leaf0;
            for (int b = T.blk0[l]; b < T.blk0[l] + T.nblk[l]; b++)
                if (!has_child[b]) c->h_leaves.push_back(b);
            T.nleaf[l] = (int)c->h_leaves.size() - c->lv[l].leaf0 - T.leafk0[l];
        }
        c->lv[l].nleaf = (int)c->h_leaves.size() - c->lv[l].leaf0;
        LevelInfo &L = c->lv[l];
        L.all_leaves = L.nblk > 0 && L.nleaf == L.nblk;
        for (int k = 0; k < L.nleaf && L.all_leaves; k++)
            if (c->h_leaves[(size_t)(L.leaf0 + k)] != L.blk0 + k) L.all_leaves = false;
    }
    // every index a kernel will follow stays inside its table
    for (size_t i = 0; i < c->h_nbr.size(); i++) {
        int v = c->h_nbr[i];
        if (!(v == NB_BOUNDARY || (v >= 0 && v < nb) || (v <= -2 && nb_slot(v) < c->nslot)))
            fail("internal: neighbour table entry out of range");
        if (v >= 0 && (c->h_level[v] != c->h_level[i / ND] || c->h_tree[v] != c->h_tree[i / ND]))
            fail("neighbour on another level or tree");
    }
    for (auto &sl : c->h_slots)
        if (sl.x < 0 || sl.x >= nb || sl.z < 0 || sl.z >= nb || sl.y < 0 || sl.y >= ND)
            fail("internal: ghost slot out of range");
    for (int b = 0; b < nb; b++) {
        int pb = c->h_parent[b];
        if (c->h_level[b] > 0 &&
            (pb < 0 || pb >= nb || c->h_level[pb] != c->h_level[b] - 1 || c->h_tree[pb] != c->h_tree[b]))
            fail("internal: parent table entry out of range");
    }
    if ((int64_t)nb * c->V >= ((int64_t)1 << 31)) fail("field too large for 32-bit cell indices");
}

// Install tables that build_tables accepted (moved out of `t`) and make the
// device side: tables, scratch, and every field reallocated. A device error
// here (out of memory) leaves the context without a tree (later calls refuse).
static void commit_tables(srmg_ctx *c, srmg_ctx &t);
static void upload_tables(srmg_ctx *c);

static void set_forest(srmg_ctx *c, const std::vector<TreeRecords> &tr) {
    srmg_ctx t;
    t.D = c->D;
    t.M = c->M;
    t.V = c->V;
    t.F = c->F;
    build_tables(&t, tr);
    commit_tables(c, t);
}

static void commit_tables(srmg_ctx *c, srmg_ctx &t) {
    c->graphs.clear();
    c->box_ok = false;
    c->nlevels = t.nlevels;
    c->nhost = t.nhost;
    c->nb = t.nb;
    c->nslot = t.nslot;
    c->lv = std::move(t.lv);
    c->trees = std::move(t.trees);
    c->groups = std::move(t.groups);
    c->h_nbr = std::move(t.h_nbr);
    c->h_coord = std::move(t.h_coord);
    c->h_parent = std::move(t.h_parent);
    c->h_sub = std::move(t.h_sub);
    c->h_level = std::move(t.h_level);
    c->h_leaves = std::move(t.h_leaves);
    c->h_tree = std::move(t.h_tree);
    c->host2dev = std::move(t.host2dev);
    c->dev2host = std::move(t.dev2host);
    c->h_slots = std::move(t.h_slots);
    try {
        upload_tables(c);
    } catch (...) {
        c->nb = 0;
        c->nlevels = 0;
        for (auto &f : c->fields)
            if (f) {
                f->data.release();
                f->alt.release();
                f->ib.release();
                f->frame.release();
            }
        throw;
    }
}

static void upload_tables(srmg_ctx *c) {
    const int D = c->D, M = c->M, ND = 2 * D, nb = c->nb, nlevels = c->nlevels, nt = (int)c->trees.size();
    // blocks with children
    std::vector<char> has_child(nb, 0);
    for (int b = 0; b < nb; b++)
        if (c->h_parent[b] >= 0) has_child[c->h_parent[b]] = 1;
    // radial weight tables (2J-2)/(2J-1), 2J/(2J-1) for J = 0..n (entry 0 unused)
    std::vector<double> w64;
    for (int l = 0; l < nlevels; l++)
        for (int d = 0; d < std::min(D, 2); d++) {
            c->lv[l].wofs[d] = (int)w64.size();
            int n = M * c->lv[l].ext[d];
            for (int J = 0; J <= n; J++) w64.push_back(J == 0 ? 0.0 : (double)(2 * J - 2) / (double)(2 * J - 1));
            for (int J = 0; J <= n; J++) w64.push_back(J == 0 ? 0.0 : (double)(2 * J) / (double)(2 * J - 1));
        }
    std::vector<float> w32(w64.size());
    for (size_t i = 0; i < w64.size(); i++) w32[i] = (float)w64[i];
    // halo tables of the tiled kernels (Halo<D>)
    const int HN = ND + ND * 2 * (D - 1);
    std::vector<int> htab((size_t)nb * HN, 0);
    for (int b = 0; b < nb; b++) {
        int *h = htab.data() + (size_t)b * HN;
        for (int f = 0; f < ND; f++) h[f] = c->h_nbr[(size_t)b * ND + f];
        for (int f = 0; f < ND; f++) {
            int e = h[f];
            if (e < 0) continue;
            int dim = f >> 1;
            for (int dd = 0; dd < D; dd++) {
                if (dd == dim) continue;
                for (int side = 0; side < 2; side++) {
                    int k = 2 * ((dd < dim) ? dd : dd - 1) + side;
                    h[ND + f * 2 * (D - 1) + k] = c->h_nbr[(size_t)e * ND + 2 * dd + side];
                }
            }
        }
    }
    std::vector<int> isleaf(nb), pix(nb);
    for (int b = 0; b < nb; b++) {
        isleaf[b] = has_child[b] ? 0 : 1;
        pix[b] = c->h_tree[b] * nlevels + c->h_level[b];
    }
    // partial-sum ranges per (tree, level): residual norms per tree
    std::vector<int2> rl((size_t)nt * nlevels), rb((size_t)nt * nlevels);
    int pl = 0, pbk = 0;
    for (int l = 0; l < nlevels; l++) {
        for (int t = 0; t < nt; t++) {
            rl[(size_t)t * nlevels + l] = make_int2(pl + c->trees[t].leafk0[l], c->trees[t].nleaf[l]);
            rb[(size_t)t * nlevels + l] = make_int2(pbk + c->trees[t].blk0[l] - c->lv[l].blk0, c->trees[t].nblk[l]);
        }
        pl += c->lv[l].nleaf;
        pbk += c->lv[l].nblk;
    }
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
        b.alloc(bytes);
        if (bytes) HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    };
    up(c->d_nbr, c->h_nbr.data(), c->h_nbr.size() * sizeof(int));
    up(c->d_coord, c->h_coord.data(), c->h_coord.size() * sizeof(int));
    up(c->d_parent, c->h_parent.data(), c->h_parent.size() * sizeof(int));
    up(c->d_sub, c->h_sub.data(), c->h_sub.size() * sizeof(int));
    up(c->d_slots, c->h_slots.data(), c->h_slots.size() * sizeof(int4));
    up(c->d_leaves, c->h_leaves.data(), c->h_leaves.size() * sizeof(int));
    up(c->d_halo, htab.data(), htab.size() * sizeof(int));
    up(c->d_isleaf, isleaf.data(), isleaf.size() * sizeof(int));
    up(c->d_pix, pix.data(), pix.size() * sizeof(int));
    up(c->d_w64, w64.data(), w64.size() * sizeof(double));
    up(c->d_w32, w32.data(), w32.size() * sizeof(float));
    up(c->d_rng_leaf, rl.data(), rl.size() * sizeof(int2));
    up(c->d_rng_blk, rb.data(), rb.size() * sizeof(int2));
    // a second processing order of the sibling groups of grouped 3D levels
    // in the column sweeps (srmg_set_group_order 1): by tree, then by slabs
    // two groups thick in z, rows of groups along y, groups along x, z within
    // the slab. A work group reads its neighbours' face layers; the x and y
    // faces of a block touch every cache line of the neighbour (only z layers
    // are contiguous), so this order keeps x and y neighbours a row or less
    // apart in time, where block (Morton) order puts some of them a level
    // octant apart. On cfg4's finest level it cuts the sweep's HBM traffic
    // from 1.16x to 1.10x of the algorithmic bytes but not its time (FP64 a
    // little slower), so block order stays the default.
    std::vector<int> gord;
    for (int l = 0; l < nlevels; l++) {
        LevelInfo &L = c->lv[l];
        L.gord0 = -1;
        if (D != 3 || !L.grouped) continue;
        const int ngr = L.nblk / 8;
        std::vector<std::array<int, 6>> key(ngr);
        for (int g = 0; g < ngr; g++) {
            const int b0 = L.blk0 + 8 * g;
            int p[3];
            for (int d = 0; d < 3; d++) p[d] = (c->h_coord[(size_t)b0 * 3 + d] - 1) >> 1;
            key[g] = {c->h_tree[b0], p[2] >> 1, p[1], p[0], p[2], g};
        }
        std::sort(key.begin(), key.end());
        L.gord0 = (int)gord.size();
        for (auto &k : key) gord.push_back(k[5]);
    }
    up(c->d_gord, gord.data(), gord.size() * sizeof(int));
    // per-block partials: at most one per block of every level
    c->maxpartials = nb + 16;
    c->d_partial.alloc((size_t)c->maxpartials * sizeof(double));
    c->d_sum.alloc(sizeof(double) * (SUM_CHUNKS + 4 + nt));
    size_t ng = 0;
    for (auto &G : c->groups) ng = std::max(ng, (size_t)G.ntree * G.N[0] * G.N[1] * (D == 3 ? G.N[2] : 1));
    c->d_grid.alloc(ng * sizeof(double));
    c->d_grid2.alloc(ng * sizeof(double));
    std::vector<std::vector<int>> btabs;
    btabs.reserve(c->groups.size());
    for (auto &G : c->groups) {
        G.cs.alloc(G.ntree * sizeof(double));
        G.shift.alloc(G.ntree * sizeof(double));
        std::vector<int> tab((size_t)G.ntree * G.nper, 0);
        for (int lb = 0; lb < G.ntree * G.nper; lb++) {
            const int blk = G.l1_blk0 + lb;
            int pos = 0;
            for (int d = D - 1; d >= 0; d--) pos = pos * (G.N[d] / M) + c->h_coord[(size_t)blk * D + d] - 1;
            tab[(size_t)(lb / G.nper) * G.nper + pos] = blk;
        }
        btabs.push_back(std::move(tab));
        up(G.btab, btabs.back().data(), btabs.back().size() * sizeof(int));
    }
    // until autotuned or set: the column sweep on two threads per column for
    // 8^3 blocks (levels without complete sibling groups run shape 0)
    c->sweep_cfg.assign(2 * nlevels, (c->D == 3 && c->M == 8) ? 8 : 0);
    c->zero[0].reset();
    c->zero[1].reset();
    c->ps.clear();
    c->pops.clear();
    c->pp_host.clear();
    c->d_pp64.alloc((size_t)3 * nt * nlevels * sizeof(double));
    c->d_pp32.alloc((size_t)3 * nt * nlevels * sizeof(float));
    for (auto &f : c->fields)
        if (f) alloc_field(c, *f);
    HIPCHK(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// per-level kernel arguments
// ---------------------------------------------------------------------------

// s_l, stencil centre (diagelm) and -(omega / diagelm) of one problem at level l
static void level_scalars(int D, int l, double s, const srmg_op *op, double omega, double &sl, double &cc,
                          double &nwc) {
    int64_t f = (int64_t)1 << l;
    sl = s / (double)(f * f);
    cc = -2.0 * D;
    if (op->disc == SRMG_HELMHOLTZ) cc = cc - sl * op->k2;
    nwc = -(omega / cc);
}

// Install the problems of the next calls: scale s[t] and operator ops[t] of
// every tree (one value each when `same`). Geometry, boundary table and
// coarsest-level treatment must agree across the trees. Uploads the
// per-(tree, level) parameter tables and the coarse-solve factors when they
// changed; must not run while a graph is being captured with new values.
static void prepare(srmg_ctx *c, const double *s, const srmg_op *ops, double omega, bool same) {
    const int nt = (int)c->trees.size();
    std::vector<double> ps(nt);
    std::vector<srmg_op> pops(nt);
    for (int t = 0; t < nt; t++) {
        ps[t] = same ? s[0] : s[t];
        pops[t] = same ? ops[0] : ops[t];
        if (pops[t].geometry != pops[0].geometry || std::memcmp(pops[t].bc, pops[0].bc, 27) != 0 ||
            pops[t].coarse != pops[0].coarse)
            fail("the trees of one call must share geometry, boundary conditions and coarse treatment");
        if (pops[t].disc != SRMG_LAPLACIAN && pops[t].disc != SRMG_HELMHOLTZ) fail("unknown discretization");
        if (pops[t].geometry < SRMG_CARTESIAN || pops[t].geometry > SRMG_CYL2) fail("unknown geometry");
        if (c->D == 3 && pops[t].geometry != SRMG_CARTESIAN) fail("cylindrical geometry is two-dimensional");
    }
    std::vector<double> pp((size_t)3 * nt * c->nlevels);
    for (int t = 0; t < nt; t++)
        for (int l = 0; l < c->nlevels; l++) {
            double *q = pp.data() + 3 * ((size_t)t * c->nlevels + l);
            level_scalars(c->D, l, ps[t], &pops[t], omega, q[0], q[1], q[2]);
        }
    bool changed = pp != c->pp_host || ps != c->ps;
    if (changed) {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(c->stream, &st));
        if (st != hipStreamCaptureStatusNone) fail("problem parameters changed during graph capture");
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<float> pf(pp.size());
        for (size_t i = 0; i < pp.size(); i++) pf[i] = (float)pp[i];
        HIPCHK(hipMemcpy(c->d_pp64.p, pp.data(), pp.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_pp32.p, pf.data(), pf.size() * sizeof(float), hipMemcpyHostToDevice));
        for (auto &G : c->groups) {
            std::vector<double> cs(G.ntree), sh(G.ntree);
            for (int t = 0; t < nt; t++) {
                if (&c->groups[c->trees[t].group] != &G) continue;
                cs[c->trees[t].gslot] = -ps[t];
                sh[c->trees[t].gslot] = (pops[t].disc == SRMG_HELMHOLTZ) ? -(ps[t] * pops[t].k2) : 0.0;
            }
            HIPCHK(hipMemcpy(G.cs.p, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(G.shift.p, sh.data(), sh.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        c->pp_host = pp;
    }
    c->ps = ps;
    c->pops = pops;
    c->pomega = omega;
}

template <typename T>
static LevelArgs<T> level_args(srmg_ctx *c, int l /*0-based*/, double s, const srmg_op *op, double omega) {
    LevelArgs<T> a;
    a.blk0 = c->lv[l].blk0;
    a.nblk = c->lv[l].nblk;
    double sl, cc, nwc;
    level_scalars(c->D, l, s, op, omega, sl, cc, nwc);
    a.s = (T)sl;
    a.cc = (T)cc;
    a.nwc = (T)nwc;
    for (int q = 0; q < 6; q++) a.sg[q] = T(1);
    for (int d = 0; d < c->D; d++)
        for (int side = 0; side < 2; side++) {
            int idx = 0, m = 1;
            for (int e = 0; e < c->D; e++) {
                int fv = (e == d) ? (side ? 1 : -1) : 0;
                idx += (fv + 1) * m;
                m *= 3;
            }
            a.sg[2 * d + side] = (T)op->bc[idx];
        }
    a.rdim = (op->geometry == SRMG_CYL1) ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
    const T *wbase = std::is_same<T, double>::value ? (const T *)c->d_w64.p : (const T *)c->d_w32.p;
    if (a.rdim >= 0) {
        int n = c->M * c->lv[l].ext[a.rdim];
        a.wm = wbase + c->lv[l].wofs[a.rdim];
        a.wp = a.wm + (n + 1);
    } else {
        a.wm = a.wp = wbase;
    }
    // several trees: per-block values from the tables installed by prepare()
    // (their -(omega/diagelm) is for c->pomega; residual kernels do not use it)
    if (c->trees.size() > 1) {
        a.pp = std::is_same<T, double>::value ? (const T *)c->d_pp64.p : (const T *)c->d_pp32.p;
        a.pix = c->d_pix.as<int>();
    } else {
        a.pp = nullptr;
        a.pix = nullptr;
    }
    a.gord = (c->group_order && c->lv[l].gord0 >= 0) ? c->d_gord.as<int>() + c->lv[l].gord0 : nullptr;
    return a;
}

static inline unsigned nwg(int64_t threads) { return (unsigned)((threads + WG - 1) / WG); }

// ---------------------------------------------------------------------------
// kernel launchers, dispatched on (D, M, T)
// ---------------------------------------------------------------------------

static std::vector<double> read_norms(srmg_ctx *c);

// n x n matrix of an axis transform: A (row-major), then its transpose (the
// coalesced operand of k_axis_gemm_c)
static std::vector<double> axis_matrix_both(const std::vector<double> &A, int n) {
    const size_t nn = (size_t)n * n;
    std::vector<double> both(2 * nn);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            both[(size_t)i * n + k] = A[(size_t)i * n + k];
            both[nn + (size_t)k * n + i] = A[(size_t)i * n + k];
        }
    return both;
}

// the same on the device
static void upload_axis_matrix(DevBuf &buf, const std::vector<double> &A, int n) {
    const size_t nn = (size_t)n * n;
    std::vector<double> both = axis_matrix_both(A, n);
    buf.alloc(2 * nn * sizeof(double));
    HIPCHK(hipMemcpy(buf.p, both.data(), 2 * nn * sizeof(double), hipMemcpyHostToDevice));
}

// one axis of the level-1 transforms (A from upload_axis_matrix). n <= 256:
// kernels without LDS where the stride allows it (variants 1 and 2),
// else the LDS-tiled kernels; the column tile of the tiled kernel is narrowed
// until the launch has at least 512 work groups. Every variant gives the same
// bits.
static void launch_axis_gemm(hipStream_t st, const double *A, int n, int inner, int ncol, const double *in,
                             double *out, int variant) {
    if (n > 64 && n <= 256 && variant >= 1) {
        hipLaunchKernelGGL(k_axis_gemm_full, dim3((unsigned)((ncol + AXF - 1) / AXF), (unsigned)((n + AXF - 1) / AXF)),
                           dim3(256), 0, st, A, n, inner, ncol, in, out);
        return;
    }
    // two rows or columns per wave, one where that leaves fewer than 1024 waves
    // (small grids: more waves in flight)
    if (n <= 256 && variant >= 1 && inner == 1) {
        const unsigned th = (unsigned)((n + 63) / 64 * 64);
        if ((int64_t)(ncol / 2) * (th / 64) >= 1024)
            hipLaunchKernelGGL(k_axis_gemm_c<2>, dim3((unsigned)((ncol + 1) / 2)), dim3(th), 0, st, A + (size_t)n * n, n,
                               ncol, in, out);
        else
            hipLaunchKernelGGL(k_axis_gemm_c<1>, dim3((unsigned)ncol), dim3(th), 0, st, A + (size_t)n * n, n, ncol, in,
                               out);
        return;
    }
    if (n <= 256 && variant >= 1 && inner % 64 == 0) {
        if ((int64_t)(ncol / 64) * ((n + 1) / 2) >= 1024)
            hipLaunchKernelGGL(k_axis_gemm_r<2>, dim3((unsigned)(ncol / 64), (unsigned)((n + 1) / 2)), dim3(64), 0, st, A,
                               n, inner, in, out);
        else
            hipLaunchKernelGGL(k_axis_gemm_r<1>, dim3((unsigned)(ncol / 64), (unsigned)n), dim3(64), 0, st, A, n, inner,
                               in, out);
        return;
    }
    if (n <= 64) {
        // one load phase: A and the column tile in LDS (16 columns per work group)
        hipLaunchKernelGGL(k_axis_gemm_small<16>, dim3((unsigned)((ncol + 15) / 16)), dim3(256), 0, st, A, n, inner,
                           ncol, in, out);
        return;
    }
    const unsigned gy = (unsigned)((n + 63) / 64);
    auto groups = [&](int tc) { return (int64_t)((ncol + tc - 1) / tc) * gy; };
    if (groups(64) >= 512)
        hipLaunchKernelGGL(k_axis_gemm<64>, dim3((unsigned)((ncol + 63) / 64), gy), dim3(256), 0, st, A, n, inner,
                           ncol, in, out);
    else if (groups(32) >= 512)
        hipLaunchKernelGGL(k_axis_gemm<32>, dim3((unsigned)((ncol + 31) / 32), gy), dim3(256), 0, st, A, n, inner,
                           ncol, in, out);
    else
        hipLaunchKernelGGL(k_axis_gemm<16>, dim3((unsigned)((ncol + 15) / 16), gy), dim3(256), 0, st, A, n, inner,
                           ncol, in, out);
}

// the three-launch level-1 solve (k_l1_fwd_xy, k_l1_z, k_l1_inv_yx) applies
// to 3D grids whose extents are multiples of 8 and at most 64
static bool l1_fused_ok(int D, const int *N) {
    if (D != 3) return false;
    for (int d = 0; d < 3; d++)
        if (N[d] > 64 || N[d] % 8) return false;
    return true;
}

// Q[d], P[d]: forward and inverse matrices of axis d as upload_axis_matrix
// lays them out (A, then its transpose); w1, w2: two grids of scratch. BLK:
// from the block field src (through btab, times cs) to the block field dst;
// otherwise from the grid gin to the grid gout.
template <int M, typename T, bool BLK>
static void launch_l1_fused(hipStream_t st, int ngrid, const int *N, const T *src, const double *gin,
                            const int *btab, const double *cs, const double *const *Q, const double *const *P,
                            const double *const *lam, const double *shift, double *w1, double *w2, T *dst,
                            double *gout) {
    constexpr int IC = 8, CC = 16, JC = 8;
    const int N0 = N[0], N1 = N[1], N2 = N[2];
    const size_t n00 = (size_t)N0 * N0, n11 = (size_t)N1 * N1, n22 = (size_t)N2 * N2;
    hipLaunchKernelGGL((k_l1_fwd_xy<M, T, BLK, IC>), dim3((unsigned)(ngrid * N2), (unsigned)(N0 / IC)), dim3(L1T), 0,
                       st, w1, src, gin, btab, cs, Q[0] + n00, Q[1] + n11, N0, N1, N2);
    hipLaunchKernelGGL((k_l1_z<CC>), dim3((unsigned)(ngrid * (N0 * N1 / CC))), dim3(L1T), 0, st, w2, w1, Q[2] + n22,
                       P[2] + n22, lam[0], lam[1], lam[2], shift, N0, N1, N2);
    hipLaunchKernelGGL((k_l1_inv_yx<M, T, BLK, JC>), dim3((unsigned)(ngrid * N2), (unsigned)(N1 / JC)), dim3(L1T), 0,
                       st, dst, gout, w2, btab, P[1] + n11, P[0] + n00, N0, N1, N2);
}

template <int D, int M, typename T>
struct Ops {
    using LY = Lay<D, M>;
    srmg_ctx *c;
    explicit Ops(srmg_ctx *ctx) : c(ctx) {}
    // values of level l (ping-pong aware); blocks of every level keep their
    // global device index in both buffers
    T *X(FieldBuf &f, int l) { return (l >= 0 && f.cur[l]) ? f.alt.as<T>() : f.data.as<T>(); }
    T *I(FieldBuf &f) { return f.ib.as<T>(); }
    const int *nbr() { return c->d_nbr.as<int>(); }
    const int *coord() { return c->d_coord.as<int>(); }

    // run fn with the radial dimension as a compile-time constant
    template <class Fn>
    void with_rd(int rdim, Fn &&fn) {
        if (rdim < 0) {
            fn(std::integral_constant<int, -1>());
            return;
        }
        if constexpr (D == 2) {
            if (rdim == 0) {
                fn(std::integral_constant<int, 0>());
                return;
            }
            if (rdim == 1) {
                fn(std::integral_constant<int, 1>());
                return;
            }
        }
        fail("cylindrical geometry is two-dimensional");
    }

    // bring every level back to the primary buffer
    void settle(FieldBuf &f) {
        for (int l = 0; l < c->nlevels; l++) {
            if (!f.cur[l]) continue;
            const LevelInfo &L = c->lv[l];
            size_t n = (size_t)L.nblk * LY::V, o = (size_t)L.blk0 * LY::V;
            if (n)
                hipLaunchKernelGGL((k_copy<T>), dim3(nwg(n)), dim3(WG), 0, c->stream, f.data.as<T>() + o,
                                   f.alt.as<T>() + o, n);
            f.cur[l] = 0;
        }
    }

    // applydelta!'s transfer of one field from the old layout (od, oib, onbr)
    // into the new one: kept blocks (cp), new blocks interpolated from their
    // old parents (ip), removed blocks restricted into their parents (rp),
    // restrictions after the copies
    void regrid_transfer(FieldBuf &f, const DevBuf &od, const DevBuf &oib, const DevBuf &onbr, const DevBuf &cp,
                         int ncp, const DevBuf &ip, int nip, const DevBuf &rp, int nrp, const LevelArgs<T> &a) {
        T *un = f.data.as<T>();
        const T *uo = od.as<T>();
        if (ncp)
            hipLaunchKernelGGL((k_regrid_copy<D, M, T>), dim3(nwg((int64_t)ncp * LY::V)), dim3(WG), 0, c->stream, un,
                               uo, cp.as<int2>(), ncp);
        if (nip)
            hipLaunchKernelGGL((k_regrid_interp<D, M, T>), dim3(nwg((int64_t)nip * LY::V)), dim3(WG), 0, c->stream,
                               un, uo, oib.as<T>(), onbr.as<int>(), ip.as<int4>(), nip, a);
        if (nrp)
            hipLaunchKernelGGL((k_regrid_restrict<D, M, T>), dim3(nwg((int64_t)nrp * (LY::V / LY::NC))), dim3(WG), 0,
                               c->stream, un, uo, rp.as<int4>(), nrp);
    }

    void gs_pass(FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l, int pass) {
        if (!a.nblk) return;
        hipLaunchKernelGGL((k_gs_pass<D, M, T>), dim3(nwg((int64_t)a.nblk * LY::H)), dim3(WG), 0, c->stream,
                           X(u, l), X(b, l), I(u), nbr(), coord(), a, pass);
    }
    // Tile shapes of the fused sweep, k = 0..3: blocks per work group x
    // work-group size. Small blocks (one colour <= 256 cells) group 1, 2, 2, 4
    // times the blocks that fill 256 threads on 256, 256, 512, 512 threads;
    // large blocks (3D, M = 16) use one block on 256, 512, 1024 or 512 threads.
    static constexpr bool SMALL = LY::H <= 256;
    static constexpr int BPW0 = SMALL ? 256 / LY::H : 1;
    // sibling-group tiles (k_gs_sweep_group) where a group's tile fits in LDS
    static constexpr bool GROUP_OK = (D == 3 && M == 8) || (D == 2 && M == 16);
    static constexpr int NCFG = GROUP_OK ? 12 : 4;
    bool grouped(int l) const { return c->group_tiles && c->lv[l].grouped; }
    template <int K> static constexpr int cfg_bpw() { return SMALL ? BPW0 * (K == 0 ? 1 : (K == 3 ? 4 : 2)) : 1; }
    template <int K> static constexpr int cfg_nt() {
        return SMALL ? (K < 2 ? 256 : 512) : (K == 0 ? 256 : (K == 2 ? 1024 : 512));
    }
    template <int K>
    void sweep_launch(T *out, T *in, FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l) {
        constexpr int BPW = cfg_bpw<K>(), NT = cfg_nt<K>();
        unsigned g = (unsigned)((a.nblk + BPW - 1) / BPW);
        with_rd(a.rdim, [&](auto rd) {
            constexpr int RD = decltype(rd)::value;
            hipLaunchKernelGGL((k_gs_sweep<D, M, T, BPW, NT, RD>), dim3(g), dim3(NT), 0, c->stream, out, in, X(b, l),
                               I(u), c->d_halo.as<int>(), coord(), a);
        });
    }
    template <int NTG, int COL = 0, int VAR = 0>
    void sweep_group_launch(T *out, T *in, FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l) {
        if constexpr (GROUP_OK && (((1 << D) * LY::H) % NTG == 0)) {
            unsigned g = (unsigned)(a.nblk >> D);
            with_rd(a.rdim, [&](auto rd) {
                constexpr int RD = decltype(rd)::value;
                // the column mappings exist for 8^3 Cartesian blocks; elsewhere
                // shapes 7 and 8 run the group kernel on NTG threads
                if constexpr (COL == 1 && D == 3 && M == 8 && RD < 0)
                    hipLaunchKernelGGL((k_gs_sweep_col<D, M, T, RD>), dim3(g), dim3(256), 0, c->stream, out, in,
                                       X(b, l), I(u), c->d_halo.as<int>(), coord(), a);
                else if constexpr (COL == 2 && D == 3 && M == 8 && RD < 0)
                    hipLaunchKernelGGL((k_gs_sweep_colz<D, M, T, RD, VAR>), dim3(g), dim3(512), 0, c->stream, out,
                                       in, X(b, l), I(u), c->d_halo.as<int>(), coord(), a);
                else
                    hipLaunchKernelGGL((k_gs_sweep_group<D, M, T, NTG, RD>), dim3(g), dim3(NTG), 0, c->stream, out,
                                       in, X(b, l), I(u), c->d_halo.as<int>(), coord(), a);
            });
        } else {
            sweep_launch<0>(out, in, u, b, a, l);
        }
    }
    // tile shapes 0..3: k_gs_sweep (cfg_bpw / cfg_nt); 4, 5, 6: sibling groups
    // on 256 / 512 / 1024 threads; 7: sibling groups, one thread per column;
    // 8: sibling groups, two threads per column; 9, 10, 11: shape 8 with nontemporal
    // stores, nontemporal right-hand-side loads, both (levels that are not grouped, and group sizes the thread
    // count does not divide, use shape 0)
    // the tile shape a sweep of level l runs
    int sweep_shape(int l, int cfg = -1) const {
        if (cfg < 0) cfg = c->sweep_cfg[(std::is_same<T, double>::value ? 0 : c->nlevels) + l];
        if (cfg >= 4 && !(GROUP_OK && c->group_tiles && c->lv[l].grouped)) cfg = 0;
        return cfg;
    }
    // the sweep of level l can take a zero input without reading it (the
    // two-threads-per-column shapes of 3D 8^3 Cartesian blocks)
    bool zin_ok(int l, int rdim) const {
        const int cfg = sweep_shape(l);
        return D == 3 && M == 8 && rdim < 0 && cfg >= 8 && cfg <= 11;
    }
    // zin: u and its ghost slabs are zero on level l (only with zin_ok(l))
    void sweep(FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l, int cfg = -1, bool zin = false) {
        if (!a.nblk) return;
        if (!u.alt.p || u.alt.bytes != u.data.bytes) u.alt.alloc(u.data.bytes);
        T *in = X(u, l);
        T *out = u.cur[l] ? u.data.as<T>() : u.alt.as<T>();
        cfg = sweep_shape(l, cfg);
        if (zin) {
            if (!zin_ok(l, a.rdim)) fail("internal: zero-input sweep on a tile shape without it");
            switch (cfg) {
                case 8: sweep_group_launch<512, 2, 4>(out, in, u, b, a, l); break;
                case 9: sweep_group_launch<512, 2, 5>(out, in, u, b, a, l); break;
                case 10: sweep_group_launch<512, 2, 6>(out, in, u, b, a, l); break;
                default: sweep_group_launch<512, 2, 7>(out, in, u, b, a, l); break;
            }
            u.cur[l] ^= 1;
            return;
        }
        switch (cfg) {
            case 1: sweep_launch<1>(out, in, u, b, a, l); break;
            case 2: sweep_launch<2>(out, in, u, b, a, l); break;
            case 3: sweep_launch<3>(out, in, u, b, a, l); break;
            case 4: sweep_group_launch<256>(out, in, u, b, a, l); break;
            case 5: sweep_group_launch<512>(out, in, u, b, a, l); break;
            case 6: sweep_group_launch<1024>(out, in, u, b, a, l); break;
            case 7: sweep_group_launch<256, 1>(out, in, u, b, a, l); break;
            case 8: sweep_group_launch<512, 2>(out, in, u, b, a, l); break;
            case 9: sweep_group_launch<512, 2, 1>(out, in, u, b, a, l); break;
            case 10: sweep_group_launch<512, 2, 2>(out, in, u, b, a, l); break;
            case 11: sweep_group_launch<512, 2, 3>(out, in, u, b, a, l); break;
            default: sweep_launch<0>(out, in, u, b, a, l); break;
        }
        u.cur[l] ^= 1;
    }
    // gauss_seidel_iter! (src/multigrid.jl:250-265)
    void smooth(FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l, int n, bool zin = false) {
        for (int i = 0; i < n; i++) sweep(u, b, a, l, -1, zin && i == 0);
    }
    // residual kernels from LDS tiles (base tile shape)
    static constexpr int RB = SMALL ? 256 / LY::H : 1;
    static constexpr int RNT = SMALL ? 256 : 512;
    // blocks per work group of k_bnorm: its per-block partials are those of
    // k_resid_tile<RB, RNT> for any count (each block's cells take the same
    // thread slots of the same halving tree, up to exact additions of zero)
    static constexpr int BB = LY::V <= 512 ? 8 : RB;
    // threads per work group of the sibling-group residual kernels: a 2D
    // group (1,024 cells) on 256 threads, eight groups resident per CU (on
    // 1,024 threads, one cell each, only two fit and the small work per
    // group left the launch latency-bound)
    static constexpr int GNT = D == 2 ? 256 : 1024;
    void residual(FieldBuf &r, FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l) {
        if (!a.nblk) return;
        if constexpr (GROUP_OK) {
            if (grouped(l)) {
                with_rd(a.rdim, [&](auto rd) {
                    constexpr int RD = decltype(rd)::value;
                    hipLaunchKernelGGL((k_resid_group<D, M, T, T, RNT, RES_WRITE, RD, GNT>),
                                       dim3((unsigned)(a.nblk >> D)), dim3(GNT), 0, c->stream, X(r, l), X(u, l), X(b, l), I(u),
                                       c->d_halo.as<int>(), coord(), (const int *)nullptr, a, 0.0,
                                       (double *)nullptr);
                });
                return;
            }
        }
        unsigned g = (unsigned)((a.nblk + RB - 1) / RB);
        with_rd(a.rdim, [&](auto rd) {
            constexpr int RD = decltype(rd)::value;
            hipLaunchKernelGGL((k_resid_tile<D, M, T, T, RB, RNT, RES_WRITE, RD>), dim3(g), dim3(RNT), 0, c->stream,
                               X(r, l), X(u, l), X(b, l), I(u), c->d_halo.as<int>(), coord(), (const int *)nullptr,
                               (const int *)nullptr, a.nblk, a, 0.0, (double *)nullptr);
        });
    }
    // residual(r, u, b, l); restrict_(u, l); [copy(u1, u, l - 1);] restrict_(r, l)
    // in one launch on grouped levels (r on level l itself is not written);
    // false when the level is not grouped (the caller runs the four steps)
    bool resid_restrict(FieldBuf &r, FieldBuf &u, FieldBuf *u1, FieldBuf &b, const LevelArgs<T> &a, int l) {
        if constexpr (GROUP_OK) {
            if (!a.nblk || !grouped(l)) return false;
            const LevelInfo &Lc = c->lv[l - 1];
            const unsigned ng = (unsigned)(a.nblk >> D);
            const int nleafc = u1 ? Lc.nleaf : 0;
            with_rd(a.rdim, [&](auto rd) {
                constexpr int RD = decltype(rd)::value;
                // 3D 8^3 Cartesian: the column form (k_resid_restrict_col)
                if constexpr (D == 3 && M == 8 && RD < 0) {
                    const unsigned ncp = (unsigned)(((int64_t)nleafc * LY::V + 255) / 256);
                    hipLaunchKernelGGL((k_resid_restrict_col<D, M, T, RD>), dim3(ng + ncp), dim3(256), 0, c->stream,
                                       X(u, l), X(b, l), I(u), c->d_halo.as<int>(), c->d_parent.as<int>(), a, ng,
                                       X(r, l - 1), X(u, l - 1), u1 ? X(*u1, l - 1) : (T *)nullptr, X(u, l - 1),
                                       c->d_leaves.as<int>() + Lc.leaf0, nleafc);
                } else {
                    const unsigned ncp = (unsigned)(((int64_t)nleafc * LY::V + GNT - 1) / GNT);
                    hipLaunchKernelGGL((k_resid_restrict<D, M, T, RNT, RD, GNT>), dim3(ng + ncp), dim3(GNT), 0,
                                       c->stream, X(u, l), X(b, l), I(u), c->d_halo.as<int>(), coord(),
                                       c->d_parent.as<int>(), a, ng, X(r, l - 1), X(u, l - 1),
                                       u1 ? X(*u1, l - 1) : (T *)nullptr, X(u, l - 1),
                                       c->d_leaves.as<int>() + Lc.leaf0, nleafc);
                }
            });
            return true;
        }
        return false;
    }
    // the per-cell gather form of the residual (kept for timing comparisons)
    void residual_gather(FieldBuf &r, FieldBuf &u, FieldBuf &b, const LevelArgs<T> &a, int l) {
        if (!a.nblk) return;
        hipLaunchKernelGGL((k_residual<D, M, T>), dim3(nwg((int64_t)a.nblk * LY::V)), dim3(WG), 0, c->stream,
                           X(r, l), X(u, l), X(b, l), I(u), nbr(), coord(), a);
    }
    // fixed-order totals of the partials: one tree -> d_sum[0] (chunked
    // two-stage sum of the first np partials); several trees -> one total per
    // tree at d_sum[TSUM + t] over its (tree, level) ranges
    void sum_partials(int np, const DevBuf &ranges) {
        if (c->trees.size() == 1) {
            int chunk = (np + SUM_CHUNKS - 1) / SUM_CHUNKS;
            double *p2 = c->d_sum.as<double>() + 4;
            hipLaunchKernelGGL(k_sum_chunks, dim3(SUM_CHUNKS), dim3(WG), 0, c->stream, c->d_partial.as<double>(), np,
                               chunk, p2);
            hipLaunchKernelGGL(k_sum, dim3(1), dim3(WG), 0, c->stream, p2, SUM_CHUNKS, c->d_sum.as<double>());
        } else {
            hipLaunchKernelGGL(k_sum_ranges, dim3((unsigned)c->trees.size()), dim3(WG), 0, c->stream,
                               c->d_partial.as<double>(), ranges.as<int2>(), c->nlevels,
                               c->d_sum.as<double>() + TSUM);
        }
    }
    void restrict_(FieldBuf &x, int l) {
        const LevelInfo &L = c->lv[l];
        if (!L.nblk) return;
        hipLaunchKernelGGL((k_restrict<D, M, T>), dim3(nwg((int64_t)L.nblk * (LY::V / LY::NC))), dim3(WG), 0,
                           c->stream, X(x, l - 1), X(x, l), c->d_parent.as<int>(), c->d_sub.as<int>(), L.blk0,
                           L.nblk);
    }
    // rhs_level! (src/multigrid.jl:378-388) of level l: b1 = R b on every block
    void rhs_level(FieldBuf &b1, FieldBuf &b, int l, const LevelArgs<T> &a, bool order4) {
        const LevelInfo &L = c->lv[l];
        if (!L.nblk) return;
        const size_t n = (size_t)L.nblk * LY::V, o = (size_t)L.blk0 * LY::V;
        if (order4)
            hipLaunchKernelGGL((k_rhs4<D, M, T>), dim3(nwg((int64_t)n)), dim3(WG), 0, c->stream, X(b1, l), X(b, l),
                               I(b), nbr(), coord(), a);
        else
            hipLaunchKernelGGL((k_copy<T>), dim3(nwg((int64_t)n)), dim3(WG), 0, c->stream, X(b1, l) + o,
                               X(b, l) + o, n);
    }
    // ---- LaplacianDiscretization{D,4}: vcycle! (src/multigrid.jl:60-106) with
    // the four-colour smoother on complete levels, step by step ----
    int rd4 = -1;  // radial dimension of the order-4 operator (set by vcycle4 / resnorm4)
    static constexpr int FP = Frame<D, M>::P;
    // the field's ghost frames (zero when first made)
    T *frame(FieldBuf &f) {
        if (!f.frame.p) {
            f.frame.alloc((size_t)c->nb * FP * sizeof(T));
            const size_t n = (size_t)c->nb * FP;
            hipLaunchKernelGGL(k_fill<T>, dim3(nwg(n)), dim3(WG), 0, c->stream, f.frame.as<T>(), (T)0, n);
        }
        return f.frame.as<T>();
    }
    // fill_ghost!(x, l) (src/connectivity.jl:155-159): copy, boundary,
    // coarse-fine interpolation; without the last step the refresh of
    // gauss_seidel_iter! (copy and boundary)
    void frame_fill(FieldBuf &x, int l, const Bc27<T> &bc, bool interp = true) {
        const LevelInfo &L = c->lv[l];
        if (!L.nblk) return;
        T *g = frame(x);
        const unsigned grid = nwg((int64_t)L.nblk * FP);
        hipLaunchKernelGGL((k_frame_copybnd<D, M, T>), dim3(grid), dim3(WG), 0, c->stream, g, X(x, l),
                           c->d_box.as<int>(), L.blk0, L.nblk, bc);
        if (interp && l > 0)
            hipLaunchKernelGGL((k_frame_interp<D, M, T>), dim3(grid), dim3(WG), 0, c->stream, g, X(x, l - 1),
                               c->d_box.as<int>(), c->d_box_slots.as<int2>(), L.blk0, L.nblk);
    }
    // copyto! / diffto! include the ghosts (src/blockfield.jl:557-576)
    void frame_copy(FieldBuf &dst, FieldBuf &src, int l) {
        const LevelInfo &L = c->lv[l];
        const size_t n = (size_t)L.nblk * FP, o = (size_t)L.blk0 * FP;
        if (n) hipLaunchKernelGGL((k_copy<T>), dim3(nwg(n)), dim3(WG), 0, c->stream, frame(dst) + o, frame(src) + o, n);
    }
    void frame_diff(FieldBuf &dst, FieldBuf &a1, FieldBuf &a2, int l) {
        const LevelInfo &L = c->lv[l];
        const size_t n = (size_t)L.nblk * FP, o = (size_t)L.blk0 * FP;
        if (n)
            hipLaunchKernelGGL((k_diff<T>), dim3(nwg(n)), dim3(WG), 0, c->stream, frame(dst) + o, frame(a1) + o,
                               frame(a2) + o, n);
    }
    void gs4_iter(FieldBuf &u, FieldBuf &b, double sl, double omega, int n, int l, const Bc27<T> &bc) {
        const LevelInfo &L = c->lv[l];
        if (!L.nblk) return;
        constexpr int NCC = D == 3 ? (M / 2) * (M / 2) * M : (M / 2) * (M / 2);
        // the reference refills the copy and boundary ghosts after every
        // colour pass; the passes after the first read those values live
        // (k_gs4_pass<LIVE>), so only the last refill is made
        for (int it = 0; it < n; it++)
            for (int col = 0; col < 4; col++) {
                const dim3 grid(nwg((int64_t)L.nblk * NCC));
                if (it == 0 && col == 0)
                    hipLaunchKernelGGL((k_gs4_pass<D, M, T, false>), grid, dim3(WG), 0, c->stream, X(u, l), X(b, l),
                                       frame(u), coord(), rd4, L.blk0, L.nblk, (T)sl, (T)omega, col >> 1, col & 1,
                                       c->d_box.as<int>(), bc);
                else
                    hipLaunchKernelGGL((k_gs4_pass<D, M, T, true>), grid, dim3(WG), 0, c->stream, X(u, l), X(b, l),
                                       frame(u), coord(), rd4, L.blk0, L.nblk, (T)sl, (T)omega, col >> 1, col & 1,
                                       c->d_box.as<int>(), bc);
            }
        if (n > 0) frame_fill(u, l, bc, false);
    }
    void vcycle4(FieldBuf &u, FieldBuf &b, FieldBuf &r, FieldBuf &u1, double s, const srmg_op *op, int lmax,
                 int ndown, int ntop, int nup, double omega) {
        Bc27<T> bc{};
        for (int k = 0; k < 27; k++) bc.v[k] = (T)op->bc[k];
        rd4 = op->geometry == SRMG_CYL1 ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
        if (lmax == 0) lmax = c->nlevels;
        for (FieldBuf *f : {&u, &b, &r, &u1}) settle(*f);
        for (int l = lmax - 1; l >= 1; l--) {
            const double sl = s / (double)((int64_t)1 << (2 * l));
            const LevelInfo &L = c->lv[l];
            gs4_iter(u, b, sl, omega, ndown, l, bc);
            restrict_(u, l);
            frame_fill(u, l - 1, bc);
            copy(u1, u, l - 1);
            frame_copy(u1, u, l - 1);
            if (L.nblk)
                hipLaunchKernelGGL((k_resid4<D, M, T>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0, c->stream,
                                   X(r, l), X(u, l), X(b, l), frame(u), coord(), rd4, L.blk0, L.nblk, (T)sl);
            restrict_(r, l);
            if (L.nblk)
                hipLaunchKernelGGL((k_postrestrict4<D, M, T>), dim3(nwg((int64_t)L.nblk * (LY::V / LY::NC))),
                                   dim3(WG), 0, c->stream, X(b, l - 1), X(u, l - 1), X(r, l - 1), frame(u),
                                   coord(), rd4, c->d_parent.as<int>(), c->d_sub.as<int>(), L.blk0, L.nblk,
                                   (T)(4 * sl));
        }
        gs4_iter(u, b, s, omega, ntop, 0, bc);
        for (int l = 1; l < lmax; l++) {
            const double sl = s / (double)((int64_t)1 << (2 * l));
            const LevelInfo &L = c->lv[l];
            diff(u1, u, u1, l - 1);
            frame_diff(u1, u, u1, l - 1);
            frame_fill(u1, l - 1, bc);
            if (L.nblk)
                hipLaunchKernelGGL((k_interp_add4<D, M, T>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0,
                                   c->stream, X(u, l), X(u1, l - 1), frame(u1), c->d_parent.as<int>(),
                                   c->d_sub.as<int>(), L.blk0, L.nblk);
            frame_fill(u, l, bc);
            gs4_iter(u, b, sl, omega, nup, l, bc);
            frame_fill(u, l, bc);
        }
    }
    // fmg! (src/multigrid.jl:17-51) with the order-4 operator
    void fmg4(FieldBuf &u, FieldBuf &b, FieldBuf &r, FieldBuf &u1, double s, const srmg_op *op, int ndown, int ntop,
              int nup, double omega) {
        Bc27<T> bc{};
        for (int k = 0; k < 27; k++) bc.v[k] = (T)op->bc[k];
        rd4 = op->geometry == SRMG_CYL1 ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
        int lmax = 0;
        for (int l = 0; l < c->nlevels; l++)
            if (c->lv[l].nblk) lmax = l + 1;
        for (FieldBuf *f : {&u, &b, &r, &u1}) settle(*f);
        for (int l = lmax - 1; l >= 1; l--) {
            const double sl = s / (double)((int64_t)1 << (2 * l));
            const LevelInfo &L = c->lv[l];
            restrict_(u, l);
            frame_fill(u, l - 1, bc);
            if (L.nblk)
                hipLaunchKernelGGL((k_resid4<D, M, T>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0, c->stream,
                                   X(r, l), X(u, l), X(b, l), frame(u), coord(), rd4, L.blk0, L.nblk, (T)sl);
            restrict_(r, l);
            if (L.nblk)
                hipLaunchKernelGGL((k_postrestrict4<D, M, T>), dim3(nwg((int64_t)L.nblk * (LY::V / LY::NC))),
                                   dim3(WG), 0, c->stream, X(b, l - 1), X(u, l - 1), X(r, l - 1), frame(u),
                                   coord(), rd4, c->d_parent.as<int>(), c->d_sub.as<int>(), L.blk0, L.nblk,
                                   (T)(4 * sl));
        }
        for (int l = 0; l < lmax; l++) {
            copy(u1, u, l);
            frame_copy(u1, u, l);
            if (l > 0) {
                const LevelInfo &L = c->lv[l];
                diff(u1, u, u1, l - 1);
                frame_diff(u1, u, u1, l - 1);
                frame_fill(u1, l - 1, bc);
                if (L.nblk)
                    hipLaunchKernelGGL((k_interp_add4<D, M, T>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0,
                                       c->stream, X(u, l), X(u1, l - 1), frame(u1), c->d_parent.as<int>(),
                                       c->d_sub.as<int>(), L.blk0, L.nblk);
            }
            frame_fill(u, l, bc);
            vcycle4(u, b, r, u1, s, op, l + 1, ndown, ntop, nup, omega);
        }
    }
    // resnorm! (src/multigrid.jl:391-417) with the order-4 operator: the
    // residual of every level into r (the reference overwrites r too), then
    // the weighted leaf sum of its squares in bnorm_partials' fixed order
    void resnorm4(FieldBuf &r, FieldBuf &u, FieldBuf &b, double s, const srmg_op *op) {
        rd4 = op->geometry == SRMG_CYL1 ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
        for (FieldBuf *f : {&u, &b, &r}) settle(*f);
        for (int l = 0; l < c->nlevels; l++) {
            const LevelInfo &L = c->lv[l];
            if (!L.nblk) continue;
            const double sl = s / (double)((int64_t)1 << (2 * l));
            hipLaunchKernelGGL((k_resid4<D, M, T>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0, c->stream,
                               X(r, l), X(u, l), X(b, l), frame(u), coord(), rd4, L.blk0, L.nblk, (T)sl);
        }
        bnorm_partials(r);
    }
    void postrestrict(FieldBuf &b, FieldBuf &u, FieldBuf &r, int l, const LevelArgs<T> &ac) {
        const LevelInfo &L = c->lv[l];
        if (!L.nblk) return;
        if (grouped(l)) {
            with_rd(ac.rdim, [&](auto rd) {
                constexpr int RD = decltype(rd)::value;
                hipLaunchKernelGGL((k_postrestrict_group<D, M, T, RD>), dim3((unsigned)(L.nblk >> D)), dim3(256), 0,
                                   c->stream, X(b, l - 1), X(u, l - 1), X(r, l - 1), I(u), c->d_halo.as<int>(),
                                   coord(), c->d_parent.as<int>(), L.blk0, ac);
            });
            return;
        }
        hipLaunchKernelGGL((k_postrestrict<D, M, T>), dim3(nwg((int64_t)L.nblk * (LY::V / LY::NC))), dim3(WG), 0,
                           c->stream, X(b, l - 1), X(u, l - 1), X(r, l - 1), I(u), nbr(), coord(),
                           c->d_parent.as<int>(), c->d_sub.as<int>(), L.blk0, L.nblk, ac);
    }
    void copy(FieldBuf &dst, FieldBuf &src, int l) {
        const LevelInfo &L = c->lv[l];
        size_t n = (size_t)L.nblk * LY::V;
        if (!n) return;
        size_t o = (size_t)L.blk0 * LY::V;
        hipLaunchKernelGGL((k_copy<T>), dim3(nwg(n)), dim3(WG), 0, c->stream, X(dst, l) + o, X(src, l) + o, n);
    }
    void diff(FieldBuf &dst, FieldBuf &a1, FieldBuf &a2, int l) {
        const LevelInfo &L = c->lv[l];
        size_t n = (size_t)L.nblk * LY::V;
        if (!n) return;
        size_t o = (size_t)L.blk0 * LY::V;
        hipLaunchKernelGGL((k_diff<T>), dim3(nwg(n)), dim3(WG), 0, c->stream, X(dst, l) + o, X(a1, l) + o,
                           X(a2, l) + o, n);
    }
    void interp_add(FieldBuf &u, FieldBuf &u1, int l, const LevelArgs<T> &ac) {
        const LevelInfo &L = c->lv[l];
        if (!L.nblk) return;
        if constexpr ((LY::V << D) % 256 == 0 && (LY::V << D) / 256 <= 16) {
            if (grouped(l)) {
                hipLaunchKernelGGL((k_interp_add_group<D, M, T, 256>), dim3((unsigned)(L.nblk >> D)), dim3(256), 0,
                                   c->stream, X(u, l), X(u1, l - 1), I(u1), c->d_halo.as<int>(),
                                   c->d_parent.as<int>(), L.blk0, ac);
                return;
            }
        }
        hipLaunchKernelGGL((k_interp_add<D, M, T>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0, c->stream,
                           X(u, l), X(u1, l - 1), I(u1), nbr(), c->d_parent.as<int>(), c->d_sub.as<int>(), L.blk0,
                           L.nblk, ac);
    }
    // diff(u1, u, u1, l) and, when acf is given, fill_interp(u1, l, *acf) in
    // one launch (the slab fill reads level l-1, the difference writes level l)
    void diff_fill(FieldBuf &u1, FieldBuf &u, int l, const LevelArgs<T> *acf) {
        const LevelInfo &L = c->lv[l];
        const size_t n = (size_t)L.nblk * LY::V;
        const int nslot = (acf && l >= 1) ? L.nslot : 0;
        if (!n && !nslot) return;
        static_assert(LY::V % (16 / sizeof(T)) == 0, "k_diff_fill takes 16 bytes per lane");
        const unsigned nd = nwg((int64_t)(n / (16 / sizeof(T)))), nf = nwg((int64_t)nslot * LY::F);
        const size_t o = (size_t)L.blk0 * LY::V;
        const LevelArgs<T> a = nslot ? *acf : LevelArgs<T>{};
        hipLaunchKernelGGL((k_diff_fill<D, M, T>), dim3(nd + nf), dim3(WG), 0, c->stream, X(u1, l) + o, X(u, l) + o,
                           X(u1, l) + o, n, nd, I(u1), X(u1, nslot ? l - 1 : l), nbr(), c->d_slots.as<int4>(),
                           L.slot0, nslot, a);
    }
    // interp_add(u, u1, l, ac) and fill_interp(u, l, ac) in one launch on
    // grouped levels (the slab fill reads level l-1 of u, final by now)
    void interp_fill(FieldBuf &u, FieldBuf &u1, int l, const LevelArgs<T> &ac) {
        const LevelInfo &L = c->lv[l];
        if constexpr ((LY::V << D) % 256 == 0 && (LY::V << D) / 256 <= 16) {
            if (L.nblk && grouped(l)) {
                const unsigned ng = (unsigned)(L.nblk >> D), nf = nwg((int64_t)L.nslot * LY::F);
                hipLaunchKernelGGL((k_interp_fill_group<D, M, T, 256>), dim3(ng + nf), dim3(256), 0, c->stream,
                                   X(u, l), X(u1, l - 1), I(u1), c->d_halo.as<int>(), c->d_parent.as<int>(), L.blk0,
                                   ng, I(u), X(u, l - 1), nbr(), c->d_slots.as<int4>(), L.slot0, L.nslot, ac);
                return;
            }
        }
        interp_add(u, u1, l, ac);
        fill_interp(u, l, ac);
    }
    // refresh the coarse-fine ghost slabs of level l from level l-1
    void fill_interp(FieldBuf &x, int l, const LevelArgs<T> &ac) {
        const LevelInfo &L = c->lv[l];
        if (!L.nslot) return;
        hipLaunchKernelGGL((k_fill_interp<D, M, T>), dim3(nwg((int64_t)L.nslot * LY::F)), dim3(WG), 0, c->stream,
                           I(x), X(x, l - 1), nbr(), c->d_slots.as<int4>(), L.slot0, L.nslot, ac);
    }
    // queue the squared-residual sums (leaf blocks, weighted by 2^(-D(l-1))) and their totals
    int resnorm_partials(FieldBuf &u, FieldBuf &b, double s, const srmg_op *op) {
        int np = 0;
        for (int l = 0; l < c->nlevels; l++) {
            const LevelInfo &L = c->lv[l];
            if (!L.nleaf) continue;
            LevelArgs<T> a = level_args<T>(c, l, s, op, 1.0);
            double wgt = std::ldexp(1.0, -c->D * l);
            unsigned g = (unsigned)((L.nleaf + RB - 1) / RB);
            with_rd(a.rdim, [&](auto rd) {
                constexpr int RD = decltype(rd)::value;
                // grouped 3D 8^3 Cartesian level of leaves in block order: the
                // column form (the same per-block sums)
                if constexpr (GROUP_OK && D == 3 && M == 8 && RD < 0 && std::is_same<T, double>::value) {
                    if (grouped(l) && L.all_leaves) {
                        hipLaunchKernelGGL((k_defect_col<D, M, RD>), dim3((unsigned)(L.nblk >> D)), dim3(256), 0,
                                           c->stream, (float *)nullptr, X(u, l), X(b, l), I(u),
                                           c->d_halo.as<int>(), c->d_isleaf.as<int>(), a, wgt,
                                           c->d_partial.as<double>() + np);
                        return;
                    }
                }
                // grouped 2D level of leaves in block order: the sibling-group
                // form, partials only (the same per-block sums)
                if constexpr (GROUP_OK && D == 2) {
                    if (grouped(l) && L.all_leaves) {
                        hipLaunchKernelGGL((k_resid_group<D, M, T, T, RNT, RES_DEFECT, RD, GNT>),
                                           dim3((unsigned)(L.nblk >> D)), dim3(GNT), 0, c->stream, (T *)nullptr,
                                           X(u, l), X(b, l), I(u), c->d_halo.as<int>(), coord(),
                                           c->d_isleaf.as<int>(), a, wgt, c->d_partial.as<double>() + np);
                        return;
                    }
                }
                hipLaunchKernelGGL((k_resid_tile<D, M, T, T, RB, RNT, RES_NORM, RD>), dim3(g), dim3(RNT), 0,
                                   c->stream, (T *)nullptr, X(u, l), X(b, l), I(u), c->d_halo.as<int>(), coord(),
                                   c->d_leaves.as<int>() + L.leaf0, (const int *)nullptr, L.nleaf, a, wgt,
                                   c->d_partial.as<double>() + np);
            });
            np += L.nleaf;
        }
        sum_partials(np, c->d_rng_leaf);
        return np;
    }
    // resnorm partials of u = 0 from b alone (bitwise those of resnorm_partials
    // on a zero field); be (may be null): (float) b on every leaf block
    int bnorm_partials(FieldBuf &b, float *be = nullptr) {
        int np = 0;
        for (int l = 0; l < c->nlevels; l++) {
            const LevelInfo &L = c->lv[l];
            if (!L.nleaf) continue;
            double wgt = std::ldexp(1.0, -c->D * l);
            unsigned g = (unsigned)((L.nleaf + BB - 1) / BB);
            hipLaunchKernelGGL((k_bnorm<D, M, T, BB, RNT>), dim3(g), dim3(RNT), 0, c->stream, X(b, l),
                               c->d_leaves.as<int>() + L.leaf0, L.nleaf, wgt, c->d_partial.as<double>() + np, be);
            np += L.nleaf;
        }
        sum_partials(np, c->d_rng_leaf);
        return np;
    }
    // u = 0 on every level and slab (values in the primary buffer)
    void zero_all(FieldBuf &u) {
        const size_t n = (size_t)c->nb * LY::V, ni = u.ib.bytes / sizeof(T);
        hipLaunchKernelGGL(k_fill<T>, dim3(nwg(n)), dim3(WG), 0, c->stream, u.data.as<T>(), (T)0, n);
        if (ni) hipLaunchKernelGGL(k_fill<T>, dim3(nwg(ni)), dim3(WG), 0, c->stream, u.ib.as<T>(), (T)0, ni);
        std::fill(u.cur.begin(), u.cur.end(), 0);
    }
    // mixed precision: be = b + L u / s (FP32) on every level, plus the leaf
    // residual norm partials, in one pass (T = double)
    void defect_and_norm(float *be, FieldBuf &u, FieldBuf &b, double s, const srmg_op *op) {
        int np = 0;
        for (int l = 0; l < c->nlevels; l++) {
            LevelArgs<T> a = level_args<T>(c, l, s, op, 1.0);
            if (!a.nblk) continue;
            double wgt = std::ldexp(1.0, -c->D * l);
            unsigned g = (unsigned)((a.nblk + RB - 1) / RB);
            bool grouped = false;
            if constexpr (GROUP_OK) grouped = this->grouped(l);
            with_rd(a.rdim, [&](auto rd) {
                constexpr int RD = decltype(rd)::value;
                if constexpr (GROUP_OK) {
                    if (grouped) {
                        // 3D 8^3 Cartesian: the column form (k_defect_col), same bits
                        if constexpr (D == 3 && M == 8 && RD < 0 && std::is_same<T, double>::value)
                            hipLaunchKernelGGL((k_defect_col<D, M, RD>), dim3((unsigned)(a.nblk >> D)), dim3(256),
                                               0, c->stream, be, X(u, l), X(b, l), I(u), c->d_halo.as<int>(),
                                               c->d_isleaf.as<int>(), a, wgt, c->d_partial.as<double>() + np);
                        else
                            hipLaunchKernelGGL((k_resid_group<D, M, T, float, RNT, RES_DEFECT, RD, GNT>),
                                               dim3((unsigned)(a.nblk >> D)), dim3(GNT), 0, c->stream, be,
                                               X(u, l), X(b, l), I(u), c->d_halo.as<int>(), coord(),
                                               c->d_isleaf.as<int>(), a, wgt, c->d_partial.as<double>() + np);
                        return;
                    }
                }
                hipLaunchKernelGGL((k_resid_tile<D, M, T, float, RB, RNT, RES_DEFECT, RD>), dim3(g), dim3(RNT), 0,
                                   c->stream, be, X(u, l), X(b, l), I(u), c->d_halo.as<int>(), coord(),
                                   (const int *)nullptr, c->d_isleaf.as<int>(), a.nblk, a, wgt,
                                   c->d_partial.as<double>() + np);
            });
            np += a.nblk;
        }
        sum_partials(np, c->d_rng_blk);
    }

    // exact FP64 solve of L_1 u + s b = 0 on the complete level 1
    void coarse_exact(FieldBuf &u, FieldBuf &b, double s, const srmg_op *op);
    // out = A in along an axis of extent n and stride `inner` (ncol columns)
    void axis(const double *A, int n, int inner, int ncol, const double *in, double *out) {
        launch_axis_gemm(c->stream, A, n, inner, ncol, in, out, c->axis_gemm);
    }

    void top(FieldBuf &u, FieldBuf &b, double s, const srmg_op *op, int ntop, double omega) {
        if (op->coarse == SRMG_COARSE_EXACT) {
            coarse_exact(u, b, s, op);
        } else {
            LevelArgs<T> a = level_args<T>(c, 0, s, op, omega);
            smooth(u, b, a, 0, ntop);
        }
    }

    // vcycle! (src/multigrid.jl:60-106). The residual of level l is formed
    // right after smoothing: it reads only level-l cells and level-l ghosts, none
    // of which the restriction to l-1 changes. The fill_ghost! calls on r are not
    // issued (r's ghosts are never read), and the descent's fill_ghost!(u, l-1)
    // needs no work: its interpolation ghosts come from level l-2, unchanged
    // since they were last filled. Levels whose sweep count is odd are copied
    // back at the end so that every cycle starts from the same buffers.
    // zin: u and its ghost slabs are zero on the finest level lmax - 1 (its
    // first sweep reads none of them; only with zin_ok)
    void vcycle(FieldBuf &u, FieldBuf &b, FieldBuf &r, FieldBuf &u1, double s, const srmg_op *op, int lmax,
                int ndown, int ntop, int nup, double omega, bool zin = false) {
        if (lmax == 0) lmax = c->nlevels;
        for (int l = lmax - 1; l >= 1; l--) {
            LevelArgs<T> a = level_args<T>(c, l, s, op, omega);
            LevelArgs<T> ac = level_args<T>(c, l - 1, s, op, omega);
            smooth(u, b, a, l, ndown, zin && l == lmax - 1);
            if (!resid_restrict(r, u, &u1, b, a, l)) {
                residual(r, u, b, a, l);
                restrict_(u, l);
                copy(u1, u, l - 1);
                restrict_(r, l);
            }
            postrestrict(b, u, r, l, ac);
        }
        top(u, b, s, op, ntop, omega);
        for (int l = 1; l < lmax; l++) {
            LevelArgs<T> a = level_args<T>(c, l, s, op, omega);
            LevelArgs<T> ac = level_args<T>(c, l - 1, s, op, omega);
            if (l - 1 >= 1) {
                LevelArgs<T> a2 = level_args<T>(c, l - 2, s, op, omega);
                diff_fill(u1, u, l - 1, &a2);
            } else {
                diff_fill(u1, u, l - 1, nullptr);
            }
            interp_fill(u, u1, l, ac);
            smooth(u, b, a, l, nup);
        }
        settle(u);
    }

    static void launch_efield(srmg_ctx *c, T *eabs, T *ef, const T *u, const T *ib, const EFieldArgs<T> &args,
                              int nblk) {
        hipLaunchKernelGGL((k_efield<D, M, T>), dim3((unsigned)nblk), dim3(256), 0, c->stream, eabs, ef, u, ib,
                           c->d_halo.as<int>(), args);
    }
    // first iterate of a solve from the level-1 problem (SRMG_GUESS_*); check:
    // its residual norm decides here whether a tree falls back to u = 0
    // (otherwise the caller does that with fall_back from a norm of its own)
    void first_iterate(FieldBuf &u, FieldBuf &b, int mode, int nsmooth, const std::vector<double> &r0,
                       bool check = true);
    int fall_back(FieldBuf &u, const std::vector<double> &rg, const std::vector<double> &r0);

    // fmg! (src/multigrid.jl:17-51)
    void fmg(FieldBuf &u, FieldBuf &b, FieldBuf &r, FieldBuf &u1, double s, const srmg_op *op, int ndown, int ntop,
             int nup, double omega) {
        int lmax = 0;
        for (int l = 0; l < c->nlevels; l++)
            if (c->lv[l].nblk) lmax = l + 1;
        for (int l = lmax - 1; l >= 1; l--) {
            LevelArgs<T> a = level_args<T>(c, l, s, op, omega);
            LevelArgs<T> ac = level_args<T>(c, l - 1, s, op, omega);
            residual(r, u, b, a, l);
            restrict_(u, l);
            restrict_(r, l);
            postrestrict(b, u, r, l, ac);
        }
        for (int l = 0; l < lmax; l++) {
            copy(u1, u, l);
            if (l > 0) {
                LevelArgs<T> ac = level_args<T>(c, l - 1, s, op, omega);
                if (l - 1 >= 1) {
                    LevelArgs<T> a2 = level_args<T>(c, l - 2, s, op, omega);
                    diff_fill(u1, u, l - 1, &a2);
                } else {
                    diff_fill(u1, u, l - 1, nullptr);
                }
                interp_fill(u, u1, l, ac);
            }
            vcycle(u, b, r, u1, s, op, l + 1, ndown, ntop, nup, omega);
        }
    }
};

// ---------------------------------------------------------------------------
// first iterate from the level-1 problem
// ---------------------------------------------------------------------------

// The charge is averaged down to level 1 (b copied into u, then restricted
// level by level), the level-1 solution of L_1 u = -s b is predicted by the
// network (LEARNED) or solved exactly (COARSE), interpolated to every finer
// level and restricted back so covered cells are the mean of their children.
// Trees whose residual is then not finite or larger than with u = 0 restart
// from zero.
template <int D, int M, typename T>
void Ops<D, M, T>::first_iterate(FieldBuf &u, FieldBuf &b, int mode, int nsmooth, const std::vector<double> &r0,
                                  bool check) {
    const double s = c->ps[0];
    const srmg_op *op = &c->pops[0];
    const int nt = (int)c->trees.size();
    if (mode == SRMG_GUESS_LEARNED) {
        if (!c->guess_net) fail("learned first iterate requested but no network is attached");
        int nd = 0, ne[3];
        srmg_guess::extents(c->guess_net, &nd, ne);
        for (const CoarseGroup &G : c->groups)
            if (nd != D || ne[0] != G.N[0] || ne[1] != G.N[1] || (D == 3 && ne[2] != G.N[2]))
                fail("the attached network was built for another level-1 grid");
        for (int t = 0; t < nt; t++) {
            const srmg_op &o = c->pops[t];
            if (!srmg_guess::matches(c->guess_net, o.geometry, o.disc, o.disc ? o.k2 : 0.0, o.bc))
                fail("the attached network was trained for another level-1 operator");
        }
    }
    settle(u);
    const size_t nall = (size_t)c->nb * LY::V;
    // b restricted to level 1: b on the coarser levels, the finest level's
    // restriction read from b itself (levels are contiguous, finest last)
    const int Lt = c->nlevels - 1;
    const size_t oL = Lt >= 1 ? (size_t)c->lv[Lt].blk0 * LY::V : nall;
    hipLaunchKernelGGL((k_copy<T>), dim3(nwg(oL)), dim3(WG), 0, c->stream, u.data.as<T>(), X(b, 0), oL);
    if (Lt >= 1 && c->lv[Lt].nblk) {
        const LevelInfo &L = c->lv[Lt];
        hipLaunchKernelGGL((k_restrict<D, M, T>), dim3(nwg((int64_t)L.nblk * (LY::V / LY::NC))), dim3(WG), 0,
                           c->stream, X(u, Lt - 1), X(b, Lt), c->d_parent.as<int>(), c->d_sub.as<int>(), L.blk0,
                           L.nblk);
    }
    for (int l = Lt - 1; l >= 1; l--) restrict_(u, l);
    // levels >= 2 and the ghost slabs start at zero
    const size_t o1 = (size_t)c->lv[0].nblk * LY::V, ni = u.ib.bytes / sizeof(T);
    if (nall > o1)
        hipLaunchKernelGGL(k_fill<T>, dim3(nwg(nall - o1)), dim3(WG), 0, c->stream, u.data.as<T>() + o1, (T)0,
                           nall - o1);
    if (ni) hipLaunchKernelGGL(k_fill<T>, dim3(nwg(ni)), dim3(WG), 0, c->stream, u.ib.as<T>(), (T)0, ni);
    if (mode == SRMG_GUESS_COARSE) {
        srmg_op oe = *op;
        oe.coarse = SRMG_COARSE_EXACT;
        coarse_exact(u, u, s, &oe);
    } else {
        for (CoarseGroup &G : c->groups) {
            const size_t total = (size_t)G.N[0] * G.N[1] * (D == 3 ? G.N[2] : 1);
            const int nblk = G.ntree * G.nper;
            double *g = c->d_grid.as<double>(), *g2 = c->d_grid2.as<double>();
            hipLaunchKernelGGL((k_grid_gather<D, M, T>), dim3(nwg((int64_t)nblk * LY::V)), dim3(WG), 0, c->stream,
                               g, X(u, 0), coord(), G.l1_blk0, nblk, G.N[0], G.N[1], G.nper, total,
                               G.cs.as<double>());
            srmg_guess::infer(c->guess_net, g, g2, G.ntree, c->stream);
            hipLaunchKernelGGL((k_grid_scatter<D, M, T>), dim3(nwg((int64_t)nblk * LY::V)), dim3(WG), 0,
                               c->stream, X(u, 0), g2, coord(), G.l1_blk0, nblk, G.N[0], G.N[1], G.nper, total);
        }
        // the prediction's grid-scale error is what the Laplacian amplifies:
        // a few level-1 sweeps remove it
        if (nsmooth > 0) {
            smooth(u, b, level_args<T>(c, 0, s, op, 1.0), 0, nsmooth);
            settle(u);
        }
    }
    for (int l = 1; l < c->nlevels; l++) interp_fill(u, u, l, level_args<T>(c, l - 1, s, op, 1.0));
    for (int l = c->nlevels - 1; l >= 1; l--) restrict_(u, l);
    for (int l = 1; l < c->nlevels; l++) fill_interp(u, l, level_args<T>(c, l - 1, s, op, 1.0));
    if (!check) return;
    resnorm_partials(u, b, s, op);
    fall_back(u, read_norms(c), r0);
}

// u = 0 on every tree whose first-iterate residual rg is not finite or exceeds
// resnorm(u = 0); returns the number of such trees (also in guess_fallbacks)
template <int D, int M, typename T>
int Ops<D, M, T>::fall_back(FieldBuf &u, const std::vector<double> &rg, const std::vector<double> &r0) {
    const double s = c->ps[0];
    const srmg_op *op = &c->pops[0];
    c->guess_fallbacks = 0;
    for (int t = 0; t < (int)c->trees.size(); t++) {
        if (std::isfinite(rg[t]) && rg[t] <= r0[t]) continue;
        c->guess_fallbacks++;
        const TreeInfo &ti = c->trees[t];
        for (int l = 0; l < c->nlevels; l++) {
            size_t n = (size_t)ti.nblk[l] * LY::V;
            if (n)
                hipLaunchKernelGGL(k_fill<T>, dim3(nwg(n)), dim3(WG), 0, c->stream,
                                   u.data.as<T>() + (size_t)ti.blk0[l] * LY::V, (T)0, n);
        }
    }
    if (c->guess_fallbacks)
        for (int l = 1; l < c->nlevels; l++) fill_interp(u, l, level_args<T>(c, l - 1, s, op, 1.0));
    return c->guess_fallbacks;
}

// ---------------------------------------------------------------------------
// exact coarsest-level solve
// ---------------------------------------------------------------------------

// Symmetric tridiagonal eigenproblem by the implicit QL method with Wilkinson
// shifts. d: diagonal (eigenvalues on return); e[i] couples i and i+1
// (destroyed); z: n x n row-major, returns eigenvectors in columns.
static bool tridiag_ql(int n, double *d, double *e, double *z) {
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) z[i * n + k] = (i == k) ? 1.0 : 0.0;
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; m++) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m == l) break;
            if (++iter > 200) return false;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, cc = 1.0, p = 0.0;
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                double f = s * e[i], bb = cc * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                cc = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * cc * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = cc * r - bb;
                for (int k = 0; k < n; k++) {
                    double zf = z[k * n + i + 1];
                    z[k * n + i + 1] = s * z[k * n + i] + cc * zf;
                    z[k * n + i] = cc * z[k * n + i] - s * zf;
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

// Level-1 1D operator along one axis (mirror ghosts, cylindrical weights on the
// radial axis) as T = P diag(lam) Q, via the symmetric form W^1/2 T W^-1/2.
static void axis_decomp(int n, bool radial, double alow, double ahigh, double *lam, double *P, double *Q) {
    std::vector<double> d(n), e(n, 0.0), z((size_t)n * n), w(n);
    for (int i = 0; i < n; i++) {
        int ii = i + 1;
        w[i] = radial ? (double)(2 * ii - 1) : 1.0;
        double diag = -2.0;
        if (!radial) {
            if (i == 0) diag += alow;
            if (i == n - 1) diag += ahigh;
        } else if (i == n - 1) {
            diag += ((double)(2 * ii) / (double)(2 * ii - 1)) * ahigh;
        }
        d[i] = diag;  // W^-1/2 (W T) W^-1/2 keeps the diagonal
        if (i + 1 < n) {
            double sij = radial ? (double)(2 * ii) : 1.0;
            e[i] = sij / std::sqrt(w[i] * (radial ? (double)(2 * ii + 1) : 1.0));
        }
    }
    if (!tridiag_ql(n, d.data(), e.data(), z.data())) fail("tridiagonal eigensolver did not converge");
    for (int k = 0; k < n; k++) lam[k] = d[k];
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            P[(size_t)i * n + k] = z[(size_t)i * n + k] / std::sqrt(w[i]);
            Q[(size_t)k * n + i] = z[(size_t)i * n + k] * std::sqrt(w[i]);
        }
}

template <int D, int M, typename T>
void Ops<D, M, T>::coarse_exact(FieldBuf &u, FieldBuf &b, double s, const srmg_op *op) {
    (void)s;  // each tree's -s and shift come from its group's tables (prepare())
    for (CoarseGroup &G : c->groups) {
        const int *N = G.N;
        bool hit = G.valid && G.geom == op->geometry && std::memcmp(G.bc, op->bc, 27) == 0;
        if (!hit) {
            hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
            HIPCHK(hipStreamIsCapturing(c->stream, &st));
            if (st != hipStreamCaptureStatusNone) fail("coarse-solve setup inside graph capture");
            c->graph_gen++;  // captured steps point to the matrices replaced here
            int rdim = (op->geometry == SRMG_CYL1) ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
            for (int d = 0; d < D; d++) {
                int il = 0, ih = 0, m = 1;
                for (int e = 0; e < D; e++) {
                    il += ((e == d ? -1 : 0) + 1) * m;
                    ih += ((e == d ? 1 : 0) + 1) * m;
                    m *= 3;
                }
                int n = N[d];
                std::vector<double> lam(n), P((size_t)n * n), Q((size_t)n * n);
                axis_decomp(n, d == rdim, (double)op->bc[il], (double)op->bc[ih], lam.data(), P.data(), Q.data());
                G.lam[d].alloc(n * sizeof(double));
                HIPCHK(hipMemcpy(G.lam[d].p, lam.data(), n * sizeof(double), hipMemcpyHostToDevice));
                upload_axis_matrix(G.P[d], P, n);
                upload_axis_matrix(G.Q[d], Q, n);
            }
            G.valid = true;
            G.geom = op->geometry;
            std::memcpy(G.bc, op->bc, 27);
        }
        const size_t total = (size_t)N[0] * N[1] * (D == 3 ? N[2] : 1);
        const int nblk = G.ntree * G.nper;
        double *g = c->d_grid.as<double>(), *g2 = c->d_grid2.as<double>();
        if (D == 3 && c->axis_gemm == 2 && l1_fused_ok(D, N)) {
            const double *Qp[3] = {G.Q[0].as<double>(), G.Q[1].as<double>(), G.Q[2].as<double>()};
            const double *Pp[3] = {G.P[0].as<double>(), G.P[1].as<double>(), G.P[2].as<double>()};
            const double *Lp[3] = {G.lam[0].as<double>(), G.lam[1].as<double>(), G.lam[2].as<double>()};
            if constexpr (D == 3)
                launch_l1_fused<M, T, true>(c->stream, G.ntree, N, X(b, 0), nullptr, G.btab.as<int>(),
                                            G.cs.as<double>(), Qp, Pp, Lp, G.shift.as<double>(), g, g2, X(u, 0),
                                            nullptr);
            continue;
        }
        hipLaunchKernelGGL((k_grid_gather<D, M, T>), dim3(nwg((int64_t)nblk * LY::V)), dim3(WG), 0, c->stream, g,
                           X(b, 0), coord(), G.l1_blk0, nblk, N[0], N[1], G.nper, total, G.cs.as<double>());
        int stride = 1;
        for (int d = 0; d < D; d++) {
            axis(G.Q[d].as<double>(), N[d], stride, (int)(G.ntree * total / N[d]), g, g2);
            std::swap(g, g2);
            stride *= N[d];
        }
        hipLaunchKernelGGL(k_eig_scale, dim3(nwg(total * G.ntree)), dim3(WG), 0, c->stream, g,
                           G.lam[0].as<double>(), G.lam[1].as<double>(),
                           D == 3 ? G.lam[2].as<double>() : G.lam[1].as<double>(), N[0], N[1], D,
                           G.shift.as<double>(), total, G.ntree);
        // inverse transforms in the reverse axis order (the last forward axis first)
        for (int d = D - 1; d >= 0; d--) {
            stride = 1;
            for (int e = 0; e < d; e++) stride *= N[e];
            axis(G.P[d].as<double>(), N[d], stride, (int)(G.ntree * total / N[d]), g, g2);
            std::swap(g, g2);
        }
        hipLaunchKernelGGL((k_grid_scatter<D, M, T>), dim3(nwg((int64_t)nblk * LY::V)), dim3(WG), 0, c->stream,
                           X(u, 0), g, coord(), G.l1_blk0, nblk, N[0], N[1], G.nper, total);
    }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------

template <class Fn>
static void dispatch(srmg_ctx *c, int dtype, Fn &&fn) {
#define SRMG_CASE(DD, MM)                                                        \
    if (c->D == DD && c->M == MM) {                                              \
        if (dtype == SRMG_F64) { Ops<DD, MM, double> o(c); fn(o); return; }      \
        if (dtype == SRMG_F32) { Ops<DD, MM, float> o(c); fn(o); return; }       \
    }
    SRMG_CASE(2, 16)
    SRMG_CASE(2, 8)
    SRMG_CASE(3, 8)
    SRMG_CASE(3, 16)
#undef SRMG_CASE
    fail("unsupported (D, M, dtype)");
}

static void check_tree(srmg_ctx *c) {
    if (c->nb == 0) fail("srmg_set_tree has not been called");
}

// Free boundary of a 2D cylindrical Poisson problem (FreeBC,
// src/freebc.jl:7-54): the kernel f of the constructor (computed by the
// caller), the two sine transforms as transposed n x n matrices, and work
// vectors; b holds the boundary values of the last srmg_freebc_apply.
struct srmg_freebc {
    int dim = 0, od = 1, level = 1;   // radial dimension (0-based), the other one
    int64_t n = 0;
    int64_t root[2] = {1, 1};
    double h = 0.0;
    DevBuf S2T, S3T, f, u1, am, b;
    bool have_b = false;
};

// the sine matrices of FFTW's RODFT10 / RODFT01 (src/freebc.jl:45-46),
// transposed: ST[j n + k] = S[k][j]
static std::vector<double> dst_matrix_t(int64_t n, int type) {
    std::vector<double> ST((size_t)(n * n));
    for (int64_t k = 0; k < n; k++)
        for (int64_t j = 0; j < n; j++) {
            double v;
            if (type == 2) v = 2.0 * std::sin(M_PI * ((double)j + 0.5) * (double)(k + 1) / (double)n);
            else v = (j == n - 1) ? ((k & 1) ? -1.0 : 1.0)
                                   : 2.0 * std::sin(M_PI * (double)(j + 1) * ((double)k + 0.5) / (double)n);
            ST[(size_t)(j * n + k)] = v;
        }
    return ST;
}

static int same_type(srmg_ctx *c, std::initializer_list<int> ids) {
    int t = -1;
    for (int id : ids) {
        int d = field(c, id).dtype;
        if (t >= 0 && d != t) fail("fields of one call must share an element type");
        t = d;
    }
    return t;
}

// ---------------------------------------------------------------------------
// host <-> device layout conversion
// ---------------------------------------------------------------------------

template <typename TH, typename TD>
static void to_device_layout(srmg_ctx *c, const TH *host, int G, std::vector<TD> &data, std::vector<TD> &ib) {
    const int D = c->D, M = c->M, S = M + 2 * G;
    const size_t SD = (D == 2) ? (size_t)S * S : (size_t)S * S * S;
    data.assign((size_t)c->nb * c->V, TD(0));
    ib.assign((size_t)std::max(c->nslot, 1) * c->F, TD(0));
    auto hoff = [&](int x, int y, int z) {  // 0-based valid coords (may be -1 / M)
        return (size_t)(x + G) + (size_t)S * ((size_t)(y + G) + (D == 3 ? (size_t)S * (size_t)(z + G) : 0));
    };
    for (int b = 0; b < c->nb; b++) {
        const TH *hb = host + (size_t)(c->dev2host[b] - 1) * SD;
        TD *db = data.data() + (size_t)b * c->V;
        int nz = (D == 3) ? M : 1;
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < M; y++)
                for (int x = 0; x < M; x++) {
                    int odd = (x + y + z + D) & 1;
                    int pos = (odd ? 0 : c->V / 2) + (x >> 1) + (M / 2) * (y + M * z);
                    db[pos] = (TD)hb[hoff(x, y, z)];
                }
    }
    // depth-1 ghost layer of coarse-fine faces
    for (int k = 0; k < c->nslot; k++) {
        int4 s = c->h_slots[k];
        int dim = s.y >> 1, dir = (s.y & 1) ? 1 : -1;
        const TH *hb = host + (size_t)(c->dev2host[s.x] - 1) * SD;
        for (int t = 0; t < c->F; t++) {
            int cc[3] = {0, 0, 0};
            int ta = t % M, tb = t / M;
            if (dim == 0) { cc[1] = ta; cc[2] = tb; }
            else if (dim == 1) { cc[0] = ta; cc[2] = tb; }
            else { cc[0] = ta; cc[1] = tb; }
            cc[dim] = dir < 0 ? -1 : M;
            ib[(size_t)k * c->F + t] = (TD)hb[hoff(cc[0], cc[1], cc[2])];
        }
    }
}

template <typename TD, typename TH>
static void from_device_layout(srmg_ctx *c, const std::vector<TD> &data, const std::vector<TD> &ib, TH *host,
                               int G, const int8_t *bcsign /* may be null: boundary ghosts untouched */) {
    const int D = c->D, M = c->M, S = M + 2 * G, V = c->V, ND = 2 * D;
    const size_t SD = (D == 2) ? (size_t)S * S : (size_t)S * S * S;
    auto hoff = [&](int x, int y, int z) {
        return (size_t)(x + G) + (size_t)S * ((size_t)(y + G) + (D == 3 ? (size_t)S * (size_t)(z + G) : 0));
    };
    auto dpos = [&](int x, int y, int z) {
        int odd = (x + y + z + D) & 1;
        return (odd ? 0 : V / 2) + (x >> 1) + (M / 2) * (y + M * z);
    };
    int nz = (D == 3) ? M : 1;
    for (int b = 0; b < c->nb; b++) {
        TH *hb = host + (size_t)(c->dev2host[b] - 1) * SD;
        const TD *db = data.data() + (size_t)b * V;
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < M; y++)
                for (int x = 0; x < M; x++) hb[hoff(x, y, z)] = (TH)db[dpos(x, y, z)];
        if (G < 1) continue;
        // depth-1 face ghosts
        for (int f = 0; f < ND; f++) {
            int dim = f >> 1, dir = (f & 1) ? 1 : -1;
            int t_ = c->h_nbr[(size_t)b * ND + f];
            if (t_ == NB_BOUNDARY && !bcsign) continue;
            for (int t = 0; t < c->F; t++) {
                int cc[3] = {0, 0, 0};
                int ta = t % M, tb = t / M;
                if (dim == 0) { cc[1] = ta; cc[2] = tb; }
                else if (dim == 1) { cc[0] = ta; cc[2] = tb; }
                else { cc[0] = ta; cc[1] = tb; }
                int g[3] = {cc[0], cc[1], cc[2]};
                g[dim] = dir < 0 ? -1 : M;
                TH v;
                if (t_ >= 0) {
                    int w[3] = {cc[0], cc[1], cc[2]};
                    w[dim] = dir < 0 ? M - 1 : 0;
                    v = (TH)data[(size_t)t_ * V + dpos(w[0], w[1], w[2])];
                } else if (t_ == NB_BOUNDARY) {
                    int w[3] = {cc[0], cc[1], cc[2]};
                    w[dim] = dir < 0 ? 0 : M - 1;
                    v = (TH)bcsign[f] * (TH)db[dpos(w[0], w[1], w[2])];
                } else {
                    v = (TH)ib[(size_t)nb_slot(t_) * c->F + t];
                }
                hb[hoff(g[0], g[1], g[2])] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// solve drivers
// ---------------------------------------------------------------------------

// Record the launches of `body` (which only enqueues kernels on c->stream)
// into gr without running them.
template <class F>
static void capture_graph(srmg_ctx *c, Graph &gr, F &&body) {
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(c->stream, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    HIPCHK(hipStreamEndCapture(c->stream, &gr.g));
    HIPCHK(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0));
    gr.gen = c->graph_gen;
}

// Run `body` either directly or via a graph captured on its first use.
template <class F>
static void run_maybe_graph(srmg_ctx *c, Graph &gr, bool use_graph, F &&body) {
    if (!use_graph) {
        body();
        return;
    }
    if (!gr.x) capture_graph(c, gr, body);
    HIPCHK(hipGraphLaunch(gr.x, c->stream));
}

// A solve's first step when it differs from the later ones (fmg! first, a
// zero start, another FP32 cycle count): run eagerly the first time, so that
// lazily built state exists, then recorded at once (a step leaves the host
// state as it found it, so recording it again changes nothing), so that the
// next solve replays it from its first step.
template <class F>
static void run_first_step(srmg_ctx *c, Graph &g, bool use_graph, F &&body) {
    const bool replay = use_graph && (g.ran || g.x);
    run_maybe_graph(c, g, replay, body);
    if (use_graph && !replay && !g.x) capture_graph(c, g, body);
    g.ran = true;
}

// The cached graph of a solve step: the key holds every value the step's
// launches bake in. Field buffers keep their addresses until a regrid or a
// field destroy, which clear the cache; coarse-solve matrices keep theirs until
// they are rebuilt, which outdates every graph captured before (dropped at the
// next solve's start). The per-tree parameter tables are read through
// pointers, so their contents may change between solves. The returned graph
// stays valid for the whole solve: nothing clears the cache while one runs.
static Graph &solve_graph(srmg_ctx *c, int kind, std::initializer_list<const FieldBuf *> fs,
                          const srmg_solve_opts *o) {
    // drop graphs captured before the last coarse-solve rebuild (no solve is
    // running here, so nothing refers to them)
    for (auto it = c->graphs.begin(); it != c->graphs.end();) {
        if (it->second->x && it->second->gen != c->graph_gen) it = c->graphs.erase(it);
        else ++it;
    }
    std::string k;
    auto add = [&](const auto &v) { k.append(reinterpret_cast<const char *>(&v), sizeof(v)); };
    add(kind);
    for (const FieldBuf *f : fs) {
        add(f->data.p);
        add(f->ib.p);
    }
    for (double v : c->ps) add(v);
    for (const srmg_op &op : c->pops) {
        add(op.geometry);
        add(op.disc);
        add(op.k2);
        k.append(reinterpret_cast<const char *>(op.bc), 27);
        add(op.coarse);
    }
    add(c->pomega);
    add(o->ndown);
    add(o->ntop);
    add(o->nup);
    add(o->omega);
    add(o->inner_cycles);
    add(o->inner_first);
    add(o->first_fmg);
    add(o->guess);
    for (int v : c->sweep_cfg) add(v);
    add(c->group_tiles);
    add(c->axis_gemm);
    add(c->group_order);
    auto &slot = c->graphs[k];
    if (!slot) slot = std::make_unique<Graph>();
    return *slot;
}

// the residual norms queued last (one per tree)
static std::vector<double> read_norms(srmg_ctx *c) {
    const int nt = (int)c->trees.size();
    std::vector<double> v(nt);
    if (nt == 1)
        HIPCHK(hipMemcpyAsync(v.data(), c->d_sum.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    else
        HIPCHK(hipMemcpyAsync(v.data(), c->d_sum.as<double>() + TSUM, nt * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &x : v) x = std::sqrt(x);
    return v;
}

static void check_opts(const srmg_solve_opts *o) {
    if (!o) fail("null options");
    if (o->max_cycles < 0 || o->ndown < 0 || o->nup < 0 || o->ntop < 0 || o->inner_cycles < 0 ||
        o->inner_first < 0)
        fail("bad solve options");
    if (o->guess < SRMG_GUESS_NONE || o->guess > SRMG_GUESS_ZERO) fail("bad first-iterate mode");
    if (o->guess_smooth < 0) fail("bad first-iterate smoothing count");
    if (o->first_fmg != 0 && o->first_fmg != 1) fail("first_fmg must be 0 or 1");
}

// relative residuals; true when every tree is at or below tol
static bool record(const std::vector<double> &r, const std::vector<double> &r0, double tol, int k, int max_cycles,
                   double *hist) {
    bool done = true;
    for (size_t t = 0; t < r.size(); t++) {
        double rel = r[t] / r0[t];
        if (hist) hist[t * (size_t)(max_cycles + 1) + k] = rel;
        if (!std::isfinite(rel)) fail("residual is not finite after cycle " + std::to_string(k));
        if (!(rel <= tol)) done = false;
    }
    return done;
}

static std::vector<double> safe_r0(std::vector<double> r0) {
    for (auto &x : r0)
        if (!(x > 0.0)) x = 1.0;
    return r0;
}

static void solve_plain(srmg_ctx *c, int ui, int bi, int ri, int u1i, const srmg_solve_opts *o, int *cycles,
                        double *hist) {
    check_opts(o);
    const double s = c->ps[0];
    const srmg_op *op = &c->pops[0];
    int t = same_type(c, {ui, bi, ri, u1i});
    FieldBuf &u = field(c, ui), &b = field(c, bi), &r = field(c, ri), &u1 = field(c, u1i);
    dispatch(c, t, [&](auto &ops) {
        using T = std::remove_reference_t<decltype(*ops.X(u, 0))>;
        bool done;
        std::vector<double> r0;
        if (o->guess == SRMG_GUESS_ZERO) {
            // u = 0: its residual is -b, so resnorm(u) = resnorm(0) from b alone
            ops.zero_all(u);
            ops.bnorm_partials(b);
            r0 = safe_r0(read_norms(c));
            done = record(r0, r0, o->tol, 0, o->max_cycles, hist);
        } else {
            for (int l = 1; l < c->nlevels; l++) ops.fill_interp(u, l, level_args<T>(c, l - 1, s, op, 1.0));
            ops.resnorm_partials(zero_field(c, t), b, s, op);
            r0 = safe_r0(read_norms(c));
            if (o->guess) ops.first_iterate(u, b, o->guess, o->guess_smooth, r0);
            ops.resnorm_partials(u, b, s, op);
            done = record(read_norms(c), r0, o->tol, 0, o->max_cycles, hist);
        }
        int k = 0;
        Graph &gr = solve_graph(c, 0, {&u, &b, &r, &u1}, o);
        Graph &gf = solve_graph(c, 2, {&u, &b, &r, &u1}, o);
        while (k < o->max_cycles && !done) {
            // a first cycle without a cached graph runs eagerly, so that lazily
            // built state (coarse-solve matrices, ping-pong buffers) exists
            // before a graph is captured; the fmg! step (first_fmg: the first
            // cycle) is captured on its second use
            const bool f = o->first_fmg && k == 0;
            Graph &g = f ? gf : gr;
            auto step = [&] {
                if (f) ops.fmg(u, b, r, u1, s, op, o->ndown, o->ntop, o->nup, o->omega);
                else ops.vcycle(u, b, r, u1, s, op, 0, o->ndown, o->ntop, o->nup, o->omega);
                ops.resnorm_partials(u, b, s, op);
            };
            if (f) {
                run_first_step(c, g, o->use_graph, step);
            } else {
                run_maybe_graph(c, g, o->use_graph && (k > 0 || g.x), step);
                g.ran = true;
            }
            k++;
            done = record(read_norms(c), r0, o->tol, k, o->max_cycles, hist);
        }
        if (cycles) *cycles = k;
    });
}

// Mixed precision defect correction (FP64 residual and update, FP32 V-cycles
// on the correction equation, FP64 coarsest-level solve).
template <int D, int M>
static void solve_mixed_dm(srmg_ctx *c, FieldBuf &u, FieldBuf &b, FieldBuf &e, FieldBuf &be, FieldBuf &re,
                           FieldBuf &ue, const srmg_solve_opts *o, int *cycles, double *hist) {
    using LY = Lay<D, M>;
    const double s = c->ps[0];
    const srmg_op *op = &c->pops[0];
    Ops<D, M, double> o64(c);
    Ops<D, M, float> o32(c);
    std::vector<double> r0;
    const bool zero_start = o->guess == SRMG_GUESS_ZERO;
    // on trees whose refined levels are all grouped, the add, the restriction
    // and the zeroing of e for the next step run as one launch per level
    bool fused = (Ops<D, M, double>::GROUP_OK) && c->group_tiles && c->nlevels > 1;
    for (int l = 1; l < c->nlevels && fused; l++) fused = c->lv[l].nblk > 0 && c->lv[l].grouped;
    // from u = 0 with the fused add: u's cells are written by the first
    // step's add (u = 0 + e on leaves, restrictions above) before anything
    // reads them, so only its ghost slabs and buffer state are cleared
    const bool zero_add = zero_start && fused;
    if (zero_start) {
        if (zero_add) {
            const size_t ni = u.ib.bytes / sizeof(double);
            if (ni) hipLaunchKernelGGL(k_fill<double>, dim3(nwg(ni)), dim3(WG), 0, c->stream, u.ib.as<double>(), 0.0, ni);
            std::fill(u.cur.begin(), u.cur.end(), 0);
        } else {
            o64.zero_all(u);
        }
        // the FP64 defect of u = 0 is b itself (b + L 0 / s rounds to b): its
        // FP32 copy on the leaf blocks comes with the norm's read of b (on
        // covered blocks the correction's right-hand side is the FAS one,
        // written before any use)
        o64.bnorm_partials(b, be.data.as<float>());
        r0 = safe_r0(read_norms(c));
    } else {
        for (int l = 1; l < c->nlevels; l++) o64.fill_interp(u, l, level_args<double>(c, l - 1, s, op, 1.0));
        o64.bnorm_partials(b);  // resnorm(u = 0), from b alone
        r0 = safe_r0(read_norms(c));
        // a first iterate is checked against r0 with the norm of the first
        // FP64 defect below, which the solve needs anyway
        if (o->guess) o64.first_iterate(u, b, o->guess, o->guess_smooth, r0, false);
    }
    auto defect_and_norm = [&] { o64.defect_and_norm(be.data.as<float>(), u, b, s, op); };
    const size_t ncell = (size_t)c->nb * LY::V, nib = e.ib.bytes / sizeof(float);
    // the first sweep of each correction solve on the finest level starts from
    // e = 0 there: with the fused add the sweep reads none of e (and the add
    // leaves e on that level as it is)
    const int Lf = c->nlevels - 1;
    const int rdim = op->geometry == SRMG_CYL1 ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
    const bool zin = fused && o32.zin_ok(Lf, rdim);
    if (fused) {
        // e = 0 (the finest level's only where fmg! reads it first)
        const size_t nclr = (zin && !o->first_fmg) ? (size_t)c->lv[Lf].blk0 * LY::V : ncell;
        if (nclr)
            hipLaunchKernelGGL(k_fill<float>, dim3(nwg(nclr)), dim3(WG), 0, c->stream, e.data.as<float>(), 0.0f, nclr);
    }
    // one outer step; first: the first step, which runs fmg! on the
    // correction equation before its V-cycles (first_fmg) and inner_first
    // V-cycles when that is set
    const int n_inner = std::max(1, (int)o->inner_cycles), n_first = o->inner_first > 0 ? o->inner_first : n_inner;
    const bool first_differs = o->first_fmg || n_first != n_inner || zero_add;
    auto correct = [&](bool first) {
        if (!fused)
            hipLaunchKernelGGL(k_fill<float>, dim3(nwg(ncell)), dim3(WG), 0, c->stream, e.data.as<float>(), 0.0f,
                               ncell);
        hipLaunchKernelGGL(k_fill<float>, dim3(nwg(nib)), dim3(WG), 0, c->stream, e.ib.as<float>(), 0.0f, nib);
        if (first && o->first_fmg) o32.fmg(e, be, re, ue, s, op, o->ndown, o->ntop, o->nup, o->omega);
        for (int k = 0; k < (first ? n_first : n_inner); k++)
            o32.vcycle(e, be, re, ue, s, op, 0, o->ndown, o->ntop, o->nup, o->omega,
                       zin && k == 0 && !(first && o->first_fmg));
        if constexpr (Ops<D, M, double>::GROUP_OK) {
            if (fused) {
            // u += e, e = 0 and the restriction of u, finest level first
            for (int l = c->nlevels - 1; l >= 1; l--) {
                const LevelInfo &L = c->lv[l], &L0 = c->lv[0];
                const unsigned ng = (unsigned)(L.nblk >> D);
                const int nleafc = (l == 1) ? L0.nleaf : 0;
                const unsigned nx = (unsigned)(((int64_t)nleafc * LY::V + 255) / 256);
                // the finest level's e stays for the next correction's
                // zero-input first sweep
                auto add = [&](auto zero, auto keep) {
                    hipLaunchKernelGGL((k_add_restrict<D, M, decltype(zero)::value, decltype(keep)::value>),
                                       dim3(ng + nx), dim3(256), 0, c->stream, u.data.as<double>(),
                                       e.data.as<float>(), c->d_isleaf.as<int>(), c->d_parent.as<int>(), L.blk0, ng,
                                       c->d_leaves.as<int>() + L0.leaf0, nleafc);
                };
                using F_ = std::false_type;
                using T_ = std::true_type;
                const bool zr = first && zero_add, kp = zin && l == Lf;
                if (zr && kp) add(T_(), T_());
                else if (zr) add(T_(), F_());
                else if (kp) add(F_(), T_());
                else add(F_(), F_());
            }
            }
        }
        if (!fused) {
            hipLaunchKernelGGL(k_add_correction, dim3(nwg(ncell)), dim3(WG), 0, c->stream, u.data.as<double>(),
                               e.data.as<float>(), ncell);
            // covered coarse cells = restriction of their children, as at the FAS
            // fixed point; otherwise the leaf residual would not pin them
            for (int l = c->nlevels - 1; l >= 1; l--) o64.restrict_(u, l);
        }
        for (int l = 1; l < c->nlevels; l++) o64.fill_interp(u, l, level_args<double>(c, l - 1, s, op, 1.0));
        defect_and_norm();
    };
    bool done;
    if (zero_start) {
        done = record(r0, r0, o->tol, 0, o->max_cycles, hist);
    } else {
        defect_and_norm();
        std::vector<double> rg = read_norms(c);
        if (o->guess && o64.fall_back(u, rg, r0)) {
            defect_and_norm();
            rg = read_norms(c);
        }
        done = record(rg, r0, o->tol, 0, o->max_cycles, hist);
    }
    int k = 0;
    Graph &gr = solve_graph(c, 1, {&u, &b, &e, &be, &re, &ue}, o);
    Graph &gf = solve_graph(c, 3, {&u, &b, &e, &be, &re, &ue}, o);
    while (k < o->max_cycles && !done) {
        const bool f = first_differs && k == 0;
        Graph &g = f ? gf : gr;
        auto step = [&] { correct(f); };
        if (f) {
            run_first_step(c, g, o->use_graph, step);
        } else {
            run_maybe_graph(c, g, o->use_graph && (k > 0 || g.x), step);
            g.ran = true;
        }
        k++;
        done = record(read_norms(c), r0, o->tol, k, o->max_cycles, hist);
    }
    if (k == 0 && zero_add) o64.zero_all(u);  // no step wrote u
    if (cycles) *cycles = k;
}

// Mixed precision with the order-4 operator: each outer step forms the FP32
// right-hand side be = b + L4 u / s on every level, runs inner_cycles FP32
// order-4 V-cycles on e from e = 0, adds e to u on every level and restricts u
// finest level first, so covered cells hold their children's mean; then the
// FP64 residual norm. No fused kernels: the steps of the order-4 path.
template <int D, int M>
static void solve_mixed4_dm(srmg_ctx *c, FieldBuf &u, FieldBuf &b, FieldBuf &e, FieldBuf &be, FieldBuf &re,
                            FieldBuf &ue, double s, const srmg_op *op, const srmg_solve_opts *o, int *cycles,
                            double *hist) {
    using LY = Lay<D, M>;
    Ops<D, M, double> o64(c);
    Ops<D, M, float> o32(c);
    const int rd = op->geometry == SRMG_CYL1 ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
    Bc27<double> b64{};
    Bc27<float> b32{};
    for (int k = 0; k < 27; k++) {
        b64.v[k] = (double)op->bc[k];
        b32.v[k] = (float)op->bc[k];
    }
    for (FieldBuf *f : {&u, &b}) o64.settle(*f);
    for (FieldBuf *f : {&e, &be, &re, &ue}) o32.settle(*f);
    FieldBuf r;  // the FP64 residual of resnorm! (scratch of this solve)
    r.dtype = SRMG_F64;
    alloc_field(c, r);
    o64.bnorm_partials(b);
    const std::vector<double> r0 = safe_r0(read_norms(c));
    bool done;
    if (o->guess == SRMG_GUESS_ZERO) {
        o64.zero_all(u);
        const size_t n = (size_t)c->nb * o64.FP;
        hipLaunchKernelGGL(k_fill<double>, dim3(nwg(n)), dim3(WG), 0, c->stream, o64.frame(u), 0.0, n);
        done = record(r0, r0, o->tol, 0, o->max_cycles, hist);
    } else {
        for (int l = 0; l < c->nlevels; l++) o64.frame_fill(u, l, b64);
        o64.resnorm4(r, u, b, s, op);
        done = record(read_norms(c), r0, o->tol, 0, o->max_cycles, hist);
    }
    const size_t ncell = (size_t)c->nb * LY::V, nfr = (size_t)c->nb * o32.FP;
    int k = 0;
    while (k < o->max_cycles && !done) {
        for (int l = 0; l < c->nlevels; l++) {
            const LevelInfo &L = c->lv[l];
            if (!L.nblk) continue;
            const double sl = s / (double)((int64_t)1 << (2 * l));
            hipLaunchKernelGGL((k_defect4<D, M>), dim3(nwg((int64_t)L.nblk * LY::V)), dim3(WG), 0, c->stream,
                               be.data.as<float>(), u.data.as<double>(), b.data.as<double>(), o64.frame(u),
                               o64.coord(), rd, L.blk0, L.nblk, sl);
        }
        o32.zero_all(e);
        hipLaunchKernelGGL(k_fill<float>, dim3(nwg(nfr)), dim3(WG), 0, c->stream, o32.frame(e), 0.f, nfr);
        const int n_inner = std::max(1, (int)o->inner_cycles);
        for (int it = 0; it < (k == 0 && o->inner_first > 0 ? o->inner_first : n_inner); it++) {
            if (it > 0)
                for (int l = 0; l < c->nlevels; l++) o32.frame_fill(e, l, b32);
            o32.vcycle4(e, be, re, ue, s, op, 0, o->ndown, o->ntop, o->nup, o->omega);
        }
        hipLaunchKernelGGL(k_add_correction, dim3(nwg(ncell)), dim3(WG), 0, c->stream, u.data.as<double>(),
                           e.data.as<float>(), ncell);
        for (int l = c->nlevels - 1; l >= 1; l--) o64.restrict_(u, l);
        for (int l = 0; l < c->nlevels; l++) o64.frame_fill(u, l, b64);
        o64.resnorm4(r, u, b, s, op);
        k++;
        done = record(read_norms(c), r0, o->tol, k, o->max_cycles, hist);
    }
    if (cycles) *cycles = k;
    HIPCHK(hipStreamSynchronize(c->stream));  // before the scratch residual is released
}

static void solve_mixed(srmg_ctx *c, int ui, int bi, int ei, int bei, int rei, int uei, const srmg_solve_opts *o,
                        int *cycles, double *hist) {
    check_opts(o);
    if (same_type(c, {ui, bi}) != SRMG_F64) fail("u and b must be FP64");
    if (same_type(c, {ei, bei, rei, uei}) != SRMG_F32) fail("correction fields must be FP32");
    FieldBuf &u = field(c, ui), &b = field(c, bi), &e = field(c, ei), &be = field(c, bei), &re = field(c, rei),
             &ue = field(c, uei);
#define SRMG_MCASE(DD, MM) \
    if (c->D == DD && c->M == MM) return solve_mixed_dm<DD, MM>(c, u, b, e, be, re, ue, o, cycles, hist);
    SRMG_MCASE(2, 16)
    SRMG_MCASE(2, 8)
    SRMG_MCASE(3, 8)
    SRMG_MCASE(3, 16)
#undef SRMG_MCASE
    fail("unsupported (D, M)");
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// device scratch of the entry point that allocates nothing
// ---------------------------------------------------------------------------

// srmg_grid_solve_noalloc keeps its transform operands and intermediate grids
// here instead of in hipMalloc'd buffers. Every GPU holds its own copy until
// the process ends; calls on one GPU share it, so calls that overlap on one
// GPU race.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// The device address of the current device's scratch array, looked up once
// per device (host code uses this address, never the array's name).
static void *scratch_address() {
    static std::mutex mu;
    static std::unordered_map<int, void *> addr;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail("hipGetSymbolAddress returned no address for the scratch array");
    addr[dev] = p;
    return p;
}

extern "C" {

const char *srmg_last_error(void) { return g_err.c_str(); }

int srmg_create(srmg_ctx **out, int D, int M) {
    return guarded([&] {
        if (!out) fail("null output pointer");
        if (!((D == 2 || D == 3) && (M == 8 || M == 16))) fail("supported: D in {2,3}, M in {8,16}");
        auto c = std::make_unique<srmg_ctx>();
        c->D = D;
        c->M = M;
        c->V = (D == 2) ? M * M : M * M * M;
        c->F = (D == 2) ? M : M * M;
        HIPCHK(hipStreamCreate(&c->stream));
        c->scratch = scratch_address();
        *out = c.release();
    });
}

int srmg_destroy(srmg_ctx *c) {
    return guarded([&] {
        if (!c) return;
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        c->fields.clear();
        hipStream_t s = c->stream;
        delete c;
        if (s) (void)hipStreamDestroy(s);
    });
}

int srmg_set_tree(srmg_ctx *c, int nlevels, int64_t nblocks, const int64_t *pairs, const int64_t *pairs_off,
                  const int64_t *neighbor, const int64_t *neighbor_off, const int64_t *boundary,
                  const int64_t *boundary_off, const int64_t *refboundary, const int64_t *refboundary_off,
                  const int64_t *child, const int64_t *child_off) {
    return guarded([&] {
        if (!c) fail("null context");
        set_forest(c, {TreeRecords{nlevels, nblocks, pairs, pairs_off, neighbor, neighbor_off, boundary, boundary_off,
                                   refboundary, refboundary_off, child, child_off}});
    });
}

int srmg_set_forest(srmg_ctx *c, int ntrees, const int *nlevels, const int64_t *nblocks,
                    const int64_t *const *pairs, const int64_t *const *pairs_off, const int64_t *const *neighbor,
                    const int64_t *const *neighbor_off, const int64_t *const *boundary,
                    const int64_t *const *boundary_off, const int64_t *const *refboundary,
                    const int64_t *const *refboundary_off, const int64_t *const *child,
                    const int64_t *const *child_off) {
    return guarded([&] {
        if (!c) fail("null context");
        if (ntrees < 1) fail("ntrees must be >= 1");
        std::vector<TreeRecords> tr;
        for (int t = 0; t < ntrees; t++)
            tr.push_back(TreeRecords{nlevels[t], nblocks[t], pairs[t], pairs_off[t], neighbor[t], neighbor_off[t],
                                     boundary[t], boundary_off[t], refboundary[t], refboundary_off[t], child[t],
                                     child_off[t]});
        set_forest(c, tr);
    });
}

// applydelta!'s field transfer (src/refinement.jl:339-370 with interp! and
// restrict! of src/streamer.jl:191-221) around set_forest: the blocks of the
// new trees are matched to the old ones by (tree, level, coordinate), so the
// caller's block ids may be reused or renumbered freely.
static void regrid_fields(srmg_ctx *c, const std::vector<TreeRecords> &tr, const int *ids, const int *modes,
                          int nf, const int8_t *bc) {
    check_tree(c);
    const int D = c->D;
    if (nf < 0 || (nf > 0 && (!ids || !modes))) fail("bad field list");
    if (tr.size() != c->trees.size()) fail("a regrid keeps the number of trees");
    bool interp = false;
    for (int k = 0; k < nf; k++) {
        field(c, ids[k]);
        if (modes[k] & ~(SRMG_REGRID_INTERP | SRMG_REGRID_RESTRICT)) fail("bad regrid mode");
        for (int j = 0; j < k; j++)
            if (ids[j] == ids[k]) fail("a field is listed twice");
        if (modes[k] & SRMG_REGRID_INTERP) interp = true;
    }
    if (interp && !bc) fail("interpolating new blocks needs the boundary conditions");
    srmg_op op{};
    if (bc) std::memcpy(op.bc, bc, 27);
    // the old values, with the coarse-fine ghosts of every level where new
    // blocks are interpolated (fill_ghost! before applydelta!, src/streamer.jl:439-445)
    for (int k = 0; k < nf; k++) {
        FieldBuf &f = field(c, ids[k]);
        dispatch(c, f.dtype, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(f, 0))>;
            o.settle(f);
            if (modes[k] & SRMG_REGRID_INTERP)
                for (int l = 1; l < c->nlevels; l++) o.fill_interp(f, l, level_args<T>(c, l - 1, 1.0, &op, 1.0));
        });
    }
    // Blocks are matched through sorted Morton codes per (tree, level): the
    // old blocks' device ranges are in that order already, the new ones are
    // sorted here as set_forest sorts them (their rank in that order is their
    // offset in the new device range), and a parent's code is its child's
    // shifted right by D, so every lookup is a merge of sorted lists.
    const int nt = (int)c->trees.size(), nlo = c->nlevels;
    int nln = 0;
    for (auto &R : tr) nln = std::max(nln, R.nlevels);
    using Key = std::pair<uint64_t, int>;  // (Morton code, old device block | rank in the new level)
    std::vector<std::vector<Key>> ol((size_t)nt * nlo), nl((size_t)nt * nln);
    for (int t = 0; t < nt; t++) {
        const TreeInfo &T = c->trees[t];
        for (int l = 0; l < nlo; l++) {
            auto &v = ol[(size_t)t * nlo + l];
            for (int b = T.blk0[l]; b < T.blk0[l] + T.nblk[l]; b++) {
                int64_t cd[3] = {1, 1, 1};
                for (int d = 0; d < D; d++) cd[d] = c->h_coord[(size_t)b * D + d];
                v.push_back({morton(cd, D), b});
            }
        }
        for (int l = 0; l < tr[t].nlevels; l++) {
            auto &v = nl[(size_t)t * nln + l];
            for (int64_t k = tr[t].pairs_off[l]; k < tr[t].pairs_off[l + 1]; k++) {
                const int64_t *pc = tr[t].pairs + k * (D + 1);
                for (int d = 0; d < D; d++)
                    if (pc[d] < 1 || pc[d] > ((int64_t)T.ext1[d] << l))
                        fail("block coordinate outside the level domain");
                v.push_back({morton(pc, D), 0});
            }
            std::sort(v.begin(), v.end());
            for (size_t i = 0; i < v.size(); i++) v[i].second = (int)i;
        }
    }
    const std::vector<Key> none;
    auto lst = [&](const std::vector<std::vector<Key>> &L, int n, int t, int l) -> const std::vector<Key> & {
        return (l < 0 || l >= n) ? none : L[(size_t)t * n + l];
    };
    // the entry of a sorted list with this code, for codes queried in
    // increasing order (cursor i only moves forward), or -1
    auto seek = [](const std::vector<Key> &v, size_t &i, uint64_t code) {
        while (i < v.size() && v[i].first < code) i++;
        return (i < v.size() && v[i].first == code) ? v[i].second : -1;
    };
    struct NewItem { int t, l, rank, src, sb; };
    std::vector<NewItem> cpn, ipn, rpn;  // kept (src = old block), new (src = old parent), removed (rank of the new parent)
    for (int t = 0; t < nt; t++)
        for (int l = 0; l < std::max(nlo, nln); l++) {
            const auto &A = lst(ol, nlo, t, l), &B = lst(nl, nln, t, l);
            const auto &Ap = lst(ol, nlo, t, l - 1), &Bp = lst(nl, nln, t, l - 1);
            size_t ia = 0, ib = 0, ca = 0, cb = 0, cpo = 0;
            while (ia < A.size() || ib < B.size()) {
                if (ib == B.size() || (ia < A.size() && A[ia].first < B[ib].first)) {
                    // removed: the parent must be kept
                    const uint64_t pc = A[ia].first >> D;
                    const int np = seek(Bp, cb, pc), op = seek(Ap, cpo, pc);
                    if (np < 0 || op < 0) fail("a removed block's parent is not kept (a regrid coarsens by one level)");
                    rpn.push_back({t, l - 1, np, A[ia].second, (int)(A[ia].first & ((1u << D) - 1))});
                    ia++;
                } else if (ia == A.size() || B[ib].first < A[ia].first) {
                    // new: the parent must have existed
                    const int op = seek(Ap, ca, B[ib].first >> D);
                    if (op < 0) fail("a new block's parent is not in the old tree (a regrid refines by one level)");
                    ipn.push_back({t, l, B[ib].second, op, (int)(B[ib].first & ((1u << D) - 1))});
                    ib++;
                } else {
                    cpn.push_back({t, l, B[ib].second, A[ia].second, 0});
                    ia++;
                    ib++;
                }
            }
        }
    // the new tables, checked before anything of the context changes: records
    // that fail leave the old tree and fields in place
    srmg_ctx staged;
    staged.D = c->D;
    staged.M = c->M;
    staged.V = c->V;
    staged.F = c->F;
    build_tables(&staged, tr);
    const std::vector<int> ocoord = c->h_coord;
    DevBuf onbr = std::move(c->d_nbr);
    std::vector<DevBuf> odata(nf), oib(nf);
    for (int k = 0; k < nf; k++) {
        FieldBuf &f = field(c, ids[k]);
        odata[k] = std::move(f.data);
        oib[k] = std::move(f.ib);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    commit_tables(c, staged);  // new tables; every field reallocated and zeroed
    auto dev = [&](const NewItem &it) {
        const int b = c->trees[it.t].blk0[it.l] + it.rank;
        if (it.rank >= c->trees[it.t].nblk[it.l]) fail("internal: regrid rank outside the level");
        return b;
    };
    std::vector<int2> cp;
    std::vector<int4> ip, rp;
    cp.reserve(cpn.size());
    for (auto &it : cpn) {
        const int b = dev(it);
        for (int d = 0; d < D; d++)
            if (c->h_coord[(size_t)b * D + d] != ocoord[(size_t)it.src * D + d]) fail("internal: regrid block mismatch");
        cp.push_back(make_int2(it.src, b));
    }
    for (auto &it : ipn) ip.push_back(make_int4(dev(it), it.src, it.sb, 0));
    for (auto &it : rpn) rp.push_back(make_int4(it.src, dev(it), it.sb, 0));
    DevBuf dcp, dip, drp;
    auto up = [&](DevBuf &d, const void *src, size_t bytes) {
        d.alloc(bytes);
        if (bytes) HIPCHK(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    };
    up(dcp, cp.data(), cp.size() * sizeof(int2));
    up(dip, ip.data(), ip.size() * sizeof(int4));
    up(drp, rp.data(), rp.size() * sizeof(int4));
    for (int k = 0; k < nf; k++) {
        FieldBuf &f = field(c, ids[k]);
        dispatch(c, f.dtype, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(f, 0))>;
            o.regrid_transfer(f, odata[k], oib[k], onbr, dcp, (int)cp.size(), dip,
                              (modes[k] & SRMG_REGRID_INTERP) ? (int)ip.size() : 0, drp,
                              (modes[k] & SRMG_REGRID_RESTRICT) ? (int)rp.size() : 0,
                              level_args<T>(c, 0, 1.0, &op, 1.0));
        });
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));  // before the old buffers are released
}

int srmg_regrid_forest(srmg_ctx *c, int ntrees, const int *nlevels, const int64_t *nblocks,
                       const int64_t *const *pairs, const int64_t *const *pairs_off,
                       const int64_t *const *neighbor, const int64_t *const *neighbor_off,
                       const int64_t *const *boundary, const int64_t *const *boundary_off,
                       const int64_t *const *refboundary, const int64_t *const *refboundary_off,
                       const int64_t *const *child, const int64_t *const *child_off, const int *fields,
                       const int *modes, int nfields, const int8_t *bc) {
    return guarded([&] {
        if (!c) fail("null context");
        if (ntrees < 1) fail("ntrees must be >= 1");
        std::vector<TreeRecords> tr;
        for (int t = 0; t < ntrees; t++)
            tr.push_back(TreeRecords{nlevels[t], nblocks[t], pairs[t], pairs_off[t], neighbor[t], neighbor_off[t],
                                     boundary[t], boundary_off[t], refboundary[t], refboundary_off[t], child[t],
                                     child_off[t]});
        regrid_fields(c, tr, fields, modes, nfields, bc);
    });
}

int srmg_regrid(srmg_ctx *c, int nlevels, int64_t nblocks, const int64_t *pairs, const int64_t *pairs_off,
                const int64_t *neighbor, const int64_t *neighbor_off, const int64_t *boundary,
                const int64_t *boundary_off, const int64_t *refboundary, const int64_t *refboundary_off,
                const int64_t *child, const int64_t *child_off, const int *fields, const int *modes, int nfields,
                const int8_t *bc) {
    return guarded([&] {
        if (!c) fail("null context");
        regrid_fields(c,
                      {TreeRecords{nlevels, nblocks, pairs, pairs_off, neighbor, neighbor_off, boundary, boundary_off,
                                   refboundary, refboundary_off, child, child_off}},
                      fields, modes, nfields, bc);
    });
}

int srmg_restrict_full(srmg_ctx *c, int id) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        dispatch(c, f.dtype, [&](auto &o) {
            o.settle(f);
            for (int l = c->nlevels - 1; l >= 1; l--) o.restrict_(f, l);
        });
        HIPCHK(hipGetLastError());
    });
}

int srmg_rhs(srmg_ctx *c, int b1, int b, const srmg_op *op) {
    return guarded([&] {
        if (!op) fail("null operator");
        check_tree(c);
        if (b1 == b) fail("rhs needs two distinct fields");
        const int t = same_type(c, {b1, b});
        if (op->disc != SRMG_LAPLACIAN && op->disc != SRMG_HELMHOLTZ && op->disc != SRMG_LAPLACIAN4)
            fail("unknown discretization");
        if (op->disc == SRMG_LAPLACIAN4 && op->geometry != SRMG_CARTESIAN && c->D != 2)
            fail("the order-4 cylindrical forms are 2D");
        FieldBuf &fb1 = field(c, b1), &fb = field(c, b);
        dispatch(c, t, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(fb, 0))>;
            o.settle(fb);
            o.settle(fb1);
            // fill_ghost!(b, l) then rhs_level!(b1, b, l), level by level
            // (src/streamer.jl:314-318)
            for (int l = 0; l < c->nlevels; l++) {
                if (l > 0) o.fill_interp(fb, l, level_args<T>(c, l - 1, 1.0, op, 1.0));
                o.rhs_level(fb1, fb, l, level_args<T>(c, l, 1.0, op, 1.0), op->disc == SRMG_LAPLACIAN4);
            }
        });
        HIPCHK(hipGetLastError());
    });
}

int srmg_field_create(srmg_ctx *c, int dtype, int *id) {
    return guarded([&] {
        if (!c || !id) fail("null argument");
        if (dtype != SRMG_F64 && dtype != SRMG_F32) fail("dtype must be SRMG_F64 or SRMG_F32");
        auto f = std::make_unique<FieldBuf>();
        f->dtype = dtype;
        if (c->nb) alloc_field(c, *f);
        c->fields.push_back(std::move(f));
        *id = (int)c->fields.size() - 1;
    });
}

int srmg_field_destroy(srmg_ctx *c, int id) {
    return guarded([&] {
        field(c, id);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->graphs.clear();
        c->fields[id].reset();
    });
}

int srmg_zero(srmg_ctx *c, int id) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        HIPCHK(hipMemsetAsync(f.data.p, 0, f.data.bytes, c->stream));
        HIPCHK(hipMemsetAsync(f.ib.p, 0, f.ib.bytes, c->stream));
        // the order-4 ghost frame too (fill!(u, 0) zeroes the ghosts)
        if (f.frame.p) HIPCHK(hipMemsetAsync(f.frame.p, 0, f.frame.bytes, c->stream));
        std::fill(f.cur.begin(), f.cur.end(), 0);
    });
}

int srmg_upload(srmg_ctx *c, int id, const void *host, int host_dtype, int G) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        if (!host) fail("null host pointer");
        if (G < 1) fail("host ghost depth G must be >= 1");
        auto go = [&](auto th) {
            using TH = decltype(th);
            const TH *h = static_cast<const TH *>(host);
            if (f.dtype == SRMG_F64) {
                std::vector<double> data, ib;
                to_device_layout(c, h, G, data, ib);
                HIPCHK(hipMemcpy(f.data.p, data.data(), data.size() * 8, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(f.ib.p, ib.data(), ib.size() * 8, hipMemcpyHostToDevice));
            } else {
                std::vector<float> data, ib;
                to_device_layout(c, h, G, data, ib);
                HIPCHK(hipMemcpy(f.data.p, data.data(), data.size() * 4, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(f.ib.p, ib.data(), ib.size() * 4, hipMemcpyHostToDevice));
            }
        };
        HIPCHK(hipStreamSynchronize(c->stream));
        std::fill(f.cur.begin(), f.cur.end(), 0);
        if (host_dtype == SRMG_F64) go(double());
        else if (host_dtype == SRMG_F32) go(float());
        else fail("bad host dtype");
    });
}

static int download_impl(srmg_ctx *c, int id, void *host, int host_dtype, int G, const int8_t *bc) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        if (!host) fail("null host pointer");
        if (G < 0) fail("bad G");
        if (!f.settled()) dispatch(c, f.dtype, [&](auto &o) { o.settle(f); });
        HIPCHK(hipStreamSynchronize(c->stream));
        int8_t sg[6] = {1, 1, 1, 1, 1, 1};
        if (bc)
            for (int d = 0; d < c->D; d++)
                for (int side = 0; side < 2; side++) {
                    int idx = 0, m = 1;
                    for (int e = 0; e < c->D; e++) {
                        idx += ((e == d ? (side ? 1 : -1) : 0) + 1) * m;
                        m *= 3;
                    }
                    sg[2 * d + side] = bc[idx];
                }
        auto go = [&](auto td) {
            using TD = decltype(td);
            std::vector<TD> data(f.data.bytes / sizeof(TD)), ib(f.ib.bytes / sizeof(TD));
            HIPCHK(hipMemcpy(data.data(), f.data.p, f.data.bytes, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(ib.data(), f.ib.p, f.ib.bytes, hipMemcpyDeviceToHost));
            if (host_dtype == SRMG_F64)
                from_device_layout(c, data, ib, static_cast<double *>(host), G, bc ? sg : nullptr);
            else
                from_device_layout(c, data, ib, static_cast<float *>(host), G, bc ? sg : nullptr);
        };
        if (host_dtype != SRMG_F64 && host_dtype != SRMG_F32) fail("bad host dtype");
        if (f.dtype == SRMG_F64) go(double());
        else go(float());
    });
}

int srmg_download(srmg_ctx *c, int id, void *host, int host_dtype, int G) {
    return download_impl(c, id, host, host_dtype, G, nullptr);
}

int srmg_download_bc(srmg_ctx *c, int id, void *host, int host_dtype, int G, const int8_t *bc) {
    return download_impl(c, id, host, host_dtype, G, bc);
}

int srmg_fill_ghost(srmg_ctx *c, int id, int level, const int8_t *bc) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        if (level < 1 || level > c->nlevels) fail("level out of range");
        if (level == 1) return;
        srmg_op op{};
        std::memcpy(op.bc, bc, 27);
        dispatch(c, f.dtype, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(f, 0))>;
            o.fill_interp(f, level - 1, level_args<T>(c, level - 2, 1.0, &op, 1.0));
        });
    });
}

int srmg_fill_ghost_all(srmg_ctx *c, int id, const int8_t *bc) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        srmg_op op{};
        std::memcpy(op.bc, bc, 27);
        dispatch(c, f.dtype, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(f, 0))>;
            for (int l = 1; l < c->nlevels; l++) o.fill_interp(f, l, level_args<T>(c, l - 1, 1.0, &op, 1.0));
        });
    });
}

static void do_vcycle(srmg_ctx *c, int u, int b, int r, int u1, int lmax, int ndown, int ntop, int nup,
                      double omega) {
    if (lmax < 0 || lmax > c->nlevels) fail("lmax out of range");
    int t = same_type(c, {u, b, r, u1});
    dispatch(c, t, [&](auto &o) {
        o.vcycle(field(c, u), field(c, b), field(c, r), field(c, u1), c->ps[0], &c->pops[0], lmax, ndown, ntop, nup,
                 omega);
    });
    HIPCHK(hipGetLastError());
}

// The box neighbour table of the order-4 operator: for every block and each
// of the 3^D - 1 directions, the same-level neighbour, the mask of the
// dimensions in which the direction leaves the domain, or a coarse-fine slot
// (the coarse block under the position and its subblock code), as
// connectivity! with the BoxStencil classifies them. Single tree.
static void box_table(srmg_ctx *c) {
    if (c->box_ok) return;
    const int D = c->D, NB3 = D == 3 ? 27 : 9;
    if (c->trees.size() != 1) fail("the order-4 operator runs on single-tree contexts");
    std::vector<std::unordered_map<uint64_t, int>> at(c->nlevels);
    for (int bk = 0; bk < c->nb; bk++) {
        int64_t q[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) q[d] = c->h_coord[(size_t)bk * D + d];
        at[c->h_level[bk]][morton(q, D)] = bk;
    }
    std::vector<int> box((size_t)c->nb * NB3, 0);
    std::vector<int2> slots;
    for (int bk = 0; bk < c->nb; bk++) {
        const int l = c->h_level[bk];
        for (int k = 0; k < NB3; k++) {
            int64_t q[3] = {1, 1, 1};
            int mask = 0;
            for (int d = 0; d < D; d++) {
                const int dir = (k / (d == 0 ? 1 : (d == 1 ? 3 : 9))) % 3 - 1;
                q[d] = c->h_coord[(size_t)bk * D + d] + dir;
                if (q[d] < 1 || q[d] > ((int64_t)c->trees[0].ext1[d] << l)) mask |= 1 << d;
            }
            int &e = box[(size_t)bk * NB3 + k];
            if (mask) {
                e = -(1 + mask);
                continue;
            }
            auto it = at[l].find(morton(q, D));
            if (it != at[l].end()) {
                e = it->second;
                continue;
            }
            int64_t pc[3] = {1, 1, 1};
            int sb = 0;
            for (int d = 0; d < D; d++) {
                pc[d] = (q[d] - 1) / 2 + 1;
                sb |= (int)((q[d] - 1) & 1) << d;
            }
            auto ip = l > 0 ? at[l - 1].find(morton(pc, D)) : at[0].end();
            if (l == 0 || ip == at[l - 1].end()) fail("proper nesting broken for the box stencil");
            e = -16 - (int)slots.size();
            slots.push_back(make_int2(ip->second, sb));
        }
    }
    c->d_box.alloc(box.size() * sizeof(int));
    HIPCHK(hipMemcpy(c->d_box.p, box.data(), box.size() * sizeof(int), hipMemcpyHostToDevice));
    c->d_box_slots.alloc(std::max<size_t>(slots.size(), 1) * sizeof(int2));
    if (!slots.empty())
        HIPCHK(hipMemcpy(c->d_box_slots.p, slots.data(), slots.size() * sizeof(int2), hipMemcpyHostToDevice));
    c->box_ok = true;
}

static void check_order4_geometry(srmg_ctx *c, const srmg_op *op) {
    if (op->geometry < SRMG_CARTESIAN || op->geometry > SRMG_CYL2) fail("unknown geometry");
    if (c->D == 3 && op->geometry != SRMG_CARTESIAN) fail("cylindrical geometry is two-dimensional");
}

static void do_vcycle4(srmg_ctx *c, int u, int b, int r, int u1, double s, const srmg_op *op, int lmax, int ndown,
                       int ntop, int nup, double omega) {
    check_order4_geometry(c, op);
    if (op->coarse != SRMG_COARSE_GS) fail("the order-4 operator uses the reference's level-1 sweeps");
    if (lmax < 0 || lmax > c->nlevels) fail("lmax out of range");
    const int t = same_type(c, {u, b, r, u1});
    if (t != SRMG_F64) fail("the order-4 operator runs in FP64");
    box_table(c);
    FieldBuf &fu = field(c, u), &fb = field(c, b), &fr = field(c, r), &fu1 = field(c, u1);
    // the cycle is launch-bound on small trees: it is captured once per
    // configuration and replayed (frames made and fields settled first, so
    // the captured launches allocate nothing and copy no ping-pong buffers)
    dispatch(c, t, [&](auto &o) {
        for (FieldBuf *f : {&fu, &fb, &fr, &fu1}) o.settle(*f);
        o.frame(fu);
        o.frame(fu1);
    });
    std::string k = "order4";
    auto add = [&](const auto &v) { k.append(reinterpret_cast<const char *>(&v), sizeof(v)); };
    for (FieldBuf *f : {&fu, &fb, &fr, &fu1}) {
        add(f->data.p);
        add(f->frame.p);
    }
    add(s);
    add(op->geometry);
    k.append(reinterpret_cast<const char *>(op->bc), 27);
    add(lmax);
    add(ndown);
    add(ntop);
    add(nup);
    add(omega);
    auto &slot = c->graphs[k];
    if (!slot) slot = std::make_unique<Graph>();
    slot->gen = c->graph_gen;
    run_maybe_graph(c, *slot, true, [&] {
        dispatch(c, t, [&](auto &o) { o.vcycle4(fu, fb, fr, fu1, s, op, lmax, ndown, ntop, nup, omega); });
    });
    HIPCHK(hipGetLastError());
}

int srmg_fill_ghost4(srmg_ctx *c, int id, const int8_t *bc) {
    return guarded([&] {
        FieldBuf &f = field(c, id);
        check_tree(c);
        if (!bc) fail("null boundary table");
        if (f.dtype != SRMG_F64) fail("the order-4 operator runs in FP64");
        box_table(c);
        dispatch(c, SRMG_F64, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(f, 0))>;
            Bc27<T> b27{};
            for (int k = 0; k < 27; k++) b27.v[k] = (T)bc[k];
            o.settle(f);
            for (int l = 0; l < c->nlevels; l++) o.frame_fill(f, l, b27);
        });
        HIPCHK(hipGetLastError());
    });
}

int srmg_vcycle(srmg_ctx *c, int u, int b, int r, int u1, double s, const srmg_op *op, int lmax, int ndown,
                int ntop, int nup, double omega) {
    return guarded([&] {
        check_tree(c);
        if (!op) fail("null op");
        if (op->disc == SRMG_LAPLACIAN4) {
            do_vcycle4(c, u, b, r, u1, s, op, lmax, ndown, ntop, nup, omega);
            return;
        }
        prepare(c, &s, op, omega, true);
        do_vcycle(c, u, b, r, u1, lmax, ndown, ntop, nup, omega);
    });
}

int srmg_vcycle_batched(srmg_ctx *c, int u, int b, int r, int u1, const double *s, const srmg_op *ops, int lmax,
                        int ndown, int ntop, int nup, double omega) {
    return guarded([&] {
        check_tree(c);
        if (!ops || !s) fail("null argument");
        prepare(c, s, ops, omega, false);
        do_vcycle(c, u, b, r, u1, lmax, ndown, ntop, nup, omega);
    });
}

int srmg_fmg(srmg_ctx *c, int u, int b, int r, int u1, double s, const srmg_op *op, int ndown, int ntop, int nup,
             double omega) {
    return guarded([&] {
        check_tree(c);
        if (!op) fail("null op");
        if (op->disc == SRMG_LAPLACIAN4) {
            check_order4_geometry(c, op);
            if (op->coarse != SRMG_COARSE_GS) fail("the order-4 operator uses the reference's level-1 sweeps");
            if (same_type(c, {u, b, r, u1}) != SRMG_F64) fail("the order-4 operator runs in FP64");
            box_table(c);
            dispatch(c, SRMG_F64, [&](auto &o) {
                o.fmg4(field(c, u), field(c, b), field(c, r), field(c, u1), s, op, ndown, ntop, nup, omega);
            });
            HIPCHK(hipGetLastError());
            return;
        }
        prepare(c, &s, op, omega, true);
        int t = same_type(c, {u, b, r, u1});
        dispatch(c, t, [&](auto &o) {
            o.fmg(field(c, u), field(c, b), field(c, r), field(c, u1), s, op, ndown, ntop, nup, omega);
        });
        HIPCHK(hipGetLastError());
    });
}

static void do_resnorm(srmg_ctx *c, int b, int u, double *out) {
    int t = same_type(c, {u, b});
    dispatch(c, t, [&](auto &o) { o.resnorm_partials(field(c, u), field(c, b), c->ps[0], &c->pops[0]); });
    std::vector<double> v = read_norms(c);
    for (size_t k = 0; k < v.size(); k++) out[k] = v[k];
}

int srmg_resnorm(srmg_ctx *c, int r, int b, int u, double s, const srmg_op *op, double *out) {
    return guarded([&] {
        check_tree(c);
        if (!op || !out) fail("null argument");
        if (c->trees.size() != 1) fail("several trees: use srmg_resnorm_batched");
        if (op->disc == SRMG_LAPLACIAN4) {
            check_order4_geometry(c, op);
            if (same_type(c, {r, b, u}) != SRMG_F64) fail("the order-4 operator runs in FP64");
            box_table(c);
            dispatch(c, SRMG_F64, [&](auto &o) { o.resnorm4(field(c, r), field(c, u), field(c, b), s, op); });
            out[0] = read_norms(c)[0];
            return;
        }
        prepare(c, &s, op, c->pomega, true);
        do_resnorm(c, b, u, out);
    });
}

int srmg_resnorm_batched(srmg_ctx *c, int r, int b, int u, const double *s, const srmg_op *ops, double *out) {
    return guarded([&] {
        check_tree(c);
        if (!ops || !s || !out) fail("null argument");
        (void)r;
        prepare(c, s, ops, c->pomega, false);
        do_resnorm(c, b, u, out);
    });
}

int srmg_synchronize(srmg_ctx *c) {
    return guarded([&] { HIPCHK(hipStreamSynchronize(c->stream)); });
}

// srmg_solve with the order-4 operator: the ghost pre-fill of u, then
// V-cycles (captured) and resnorm! until the relative residual is <= tol;
// resnorm(0) from b alone (the residual of u = 0 is -b)
static void solve4(srmg_ctx *c, int ui, int bi, int ri, int u1i, double s, const srmg_op *op,
                   const srmg_solve_opts *o, int *cycles, double *hist) {
    check_order4_geometry(c, op);
    if (op->coarse != SRMG_COARSE_GS) fail("the order-4 operator uses the reference's level-1 sweeps");
    if (o->guess != SRMG_GUESS_NONE && o->guess != SRMG_GUESS_ZERO)
        fail("the order-4 solve starts from u or from zero");
    if (o->first_fmg) fail("first_fmg is for the order-2 operators");
    if (same_type(c, {ui, bi, ri, u1i}) != SRMG_F64) fail("the order-4 operator runs in FP64");
    box_table(c);
    FieldBuf &u = field(c, ui), &b = field(c, bi), &r = field(c, ri);
    std::vector<double> r0;
    bool done = false;
    dispatch(c, SRMG_F64, [&](auto &ops) {
        using T = std::remove_reference_t<decltype(*ops.X(u, 0))>;
        ops.settle(b);
        ops.bnorm_partials(b);
        r0 = safe_r0(read_norms(c));
        if (o->guess == SRMG_GUESS_ZERO) {
            ops.zero_all(u);
            const size_t n = (size_t)c->nb * ops.FP;
            hipLaunchKernelGGL(k_fill<T>, dim3(nwg(n)), dim3(WG), 0, c->stream, ops.frame(u), (T)0, n);
            done = record(r0, r0, o->tol, 0, o->max_cycles, hist);
        } else {
            Bc27<T> b27{};
            for (int k = 0; k < 27; k++) b27.v[k] = (T)op->bc[k];
            ops.settle(u);
            for (int l = 0; l < c->nlevels; l++) ops.frame_fill(u, l, b27);
            ops.resnorm4(r, u, b, s, op);
            done = record(read_norms(c), r0, o->tol, 0, o->max_cycles, hist);
        }
    });
    int k = 0;
    while (k < o->max_cycles && !done) {
        do_vcycle4(c, ui, bi, ri, u1i, s, op, 0, o->ndown, o->ntop, o->nup, o->omega);
        dispatch(c, SRMG_F64, [&](auto &ops) { ops.resnorm4(r, u, b, s, op); });
        k++;
        done = record(read_norms(c), r0, o->tol, k, o->max_cycles, hist);
    }
    if (cycles) *cycles = k;
}

int srmg_solve(srmg_ctx *c, int u, int b, int r, int u1, double s, const srmg_op *op,
               const srmg_solve_opts *opts, int *cycles, double *hist) {
    return guarded([&] {
        check_tree(c);
        if (!op) fail("null op");
        check_opts(opts);
        if (op->disc == SRMG_LAPLACIAN4) {
            if (c->trees.size() != 1) fail("the order-4 operator runs on single-tree contexts");
            solve4(c, u, b, r, u1, s, op, opts, cycles, hist);
            return;
        }
        prepare(c, &s, op, opts->omega, true);
        solve_plain(c, u, b, r, u1, opts, cycles, hist);
    });
}

int srmg_solve_batched(srmg_ctx *c, int u, int b, int r, int u1, const double *s, const srmg_op *ops,
                       const srmg_solve_opts *opts, int *cycles, double *hist) {
    return guarded([&] {
        check_tree(c);
        if (!ops || !s) fail("null argument");
        check_opts(opts);
        prepare(c, s, ops, opts->omega, false);
        solve_plain(c, u, b, r, u1, opts, cycles, hist);
    });
}

int srmg_solve_mixed(srmg_ctx *c, int u, int b, int e, int be, int re, int ue, double s, const srmg_op *op,
                     const srmg_solve_opts *opts, int *cycles, double *hist) {
    return guarded([&] {
        check_tree(c);
        if (!op) fail("null op");
        check_opts(opts);
        if (op->disc == SRMG_LAPLACIAN4) {
            check_order4_geometry(c, op);
            if (c->trees.size() != 1) fail("the order-4 operator runs on single-tree contexts");
            if (op->coarse != SRMG_COARSE_GS) fail("the order-4 operator uses the reference's level-1 sweeps");
            if (opts->guess != SRMG_GUESS_NONE && opts->guess != SRMG_GUESS_ZERO)
                fail("the order-4 solve starts from u or from zero");
            if (opts->first_fmg) fail("first_fmg is for the order-2 operators");
            if (same_type(c, {u, b}) != SRMG_F64) fail("u and b must be FP64");
            if (same_type(c, {e, be, re, ue}) != SRMG_F32) fail("correction fields must be FP32");
            box_table(c);
            FieldBuf &fu = field(c, u), &fb = field(c, b), &fe = field(c, e), &fbe = field(c, be),
                     &fre = field(c, re), &fue = field(c, ue);
#define SRMG_M4CASE(DD, MM)                                                                            \
    if (c->D == DD && c->M == MM) {                                                                    \
        solve_mixed4_dm<DD, MM>(c, fu, fb, fe, fbe, fre, fue, s, op, opts, cycles, hist);              \
        return;                                                                                        \
    }
            SRMG_M4CASE(2, 16)
            SRMG_M4CASE(2, 8)
            SRMG_M4CASE(3, 8)
            SRMG_M4CASE(3, 16)
#undef SRMG_M4CASE
            fail("unsupported (D, M)");
        }
        prepare(c, &s, op, opts->omega, true);
        solve_mixed(c, u, b, e, be, re, ue, opts, cycles, hist);
    });
}

int srmg_solve_mixed_batched(srmg_ctx *c, int u, int b, int e, int be, int re, int ue, const double *s,
                             const srmg_op *ops, const srmg_solve_opts *opts, int *cycles, double *hist) {
    return guarded([&] {
        check_tree(c);
        if (!ops || !s) fail("null argument");
        check_opts(opts);
        prepare(c, s, ops, opts->omega, false);
        solve_mixed(c, u, b, e, be, re, ue, opts, cycles, hist);
    });
}

int srmg_tree_count(srmg_ctx *c, int *ntrees) {
    return guarded([&] {
        check_tree(c);
        if (!ntrees) fail("null argument");
        *ntrees = (int)c->trees.size();
    });
}

int srmg_time_kernel(srmg_ctx *c, int kind, int level, int u, int b, int r, double s, const srmg_op *op,
                     int reps, double *ms_per_launch) {
    return guarded([&] {
        check_tree(c);
        if (!op || !ms_per_launch || reps < 1) fail("bad arguments");
        if (level < 1 || level > c->nlevels) fail("level out of range");
        prepare(c, &s, op, 1.0, true);
        int t = same_type(c, {u, b, r});
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        float ms = 0.f;
        dispatch(c, t, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(field(c, u), 0))>;
            LevelArgs<T> a = level_args<T>(c, level - 1, s, op, 1.0);
            int l = level - 1;
            auto launch = [&](int i) {
                if (kind == SRMG_KERNEL_GS_PASS) o.gs_pass(field(c, u), field(c, b), a, l, i & 1);
                else if (kind == SRMG_KERNEL_RESIDUAL) o.residual(field(c, r), field(c, u), field(c, b), a, l);
                else if (kind == SRMG_KERNEL_RESIDUAL_GATHER)
                    o.residual_gather(field(c, r), field(c, u), field(c, b), a, l);
                else if (kind == SRMG_KERNEL_GS_SWEEP) o.sweep(field(c, u), field(c, b), a, l);
                else if (kind == SRMG_KERNEL_STREAM) {
                    // r = u + b over the level: reads 2, writes 1 value per cell
                    size_t n = (size_t)c->lv[l].nblk * c->V, off = (size_t)c->lv[l].blk0 * c->V;
                    hipLaunchKernelGGL((k_triad<T>), dim3(nwg(n)), dim3(WG), 0, c->stream, o.X(field(c, r), l) + off,
                                       o.X(field(c, u), l) + off, o.X(field(c, b), l) + off, n);
                } else if (kind == SRMG_KERNEL_STREAM16) {
                    size_t n = (size_t)c->lv[l].nblk * c->V, off = (size_t)c->lv[l].blk0 * c->V;
                    const size_t K = 16 / sizeof(T);
                    hipLaunchKernelGGL((k_triad16<T>), dim3(nwg((int64_t)(n / K))), dim3(WG), 0, c->stream,
                                       o.X(field(c, r), l) + off, o.X(field(c, u), l) + off,
                                       o.X(field(c, b), l) + off, n);
                } else fail("unknown kernel kind");
            };
            launch(0);
            launch(1);
            if (kind == SRMG_KERNEL_GS_SWEEP || kind == SRMG_KERNEL_STREAM || kind == SRMG_KERNEL_STREAM16) {
                // the launches captured in a graph and replayed, as a solve
                // runs its sweeps (timed eagerly, the plain-store column sweep
                // of cfg4's finest level took 67.8 us, inside a solve's graph
                // 62.7 us); the faster of two replays counts
                Graph gr;
                capture_graph(c, gr, [&] {
                    for (int i = 0; i < reps; i++) launch(i);
                });
                HIPCHK(hipGraphLaunch(gr.x, c->stream));
                ms = 1e30f;
                for (int trial = 0; trial < 2; trial++) {
                    HIPCHK(hipEventRecord(e0, c->stream));
                    HIPCHK(hipGraphLaunch(gr.x, c->stream));
                    HIPCHK(hipEventRecord(e1, c->stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float t = 0.f;
                    HIPCHK(hipEventElapsedTime(&t, e0, e1));
                    ms = std::min(ms, t);
                }
            } else {
                HIPCHK(hipEventRecord(e0, c->stream));
                for (int i = 0; i < reps; i++) launch(i);
                HIPCHK(hipEventRecord(e1, c->stream));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            }
            o.settle(field(c, u));
        });
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_launch = ms / reps;
    });
}

int srmg_autotune(srmg_ctx *c, int u, int b, double s, const srmg_op *op, int reps, int *chosen) {
    return guarded([&] {
        check_tree(c);
        if (!op || reps < 1) fail("bad arguments");
        prepare(c, &s, op, 1.0, true);
        int t = same_type(c, {u, b});
        FieldBuf &fu = field(c, u);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        dispatch(c, t, [&](auto &o) {
            using T = std::remove_reference_t<decltype(*o.X(fu, 0))>;
            o.settle(fu);
            // the trial sweeps run on a copy of u
            DevBuf keep;
            keep.alloc(fu.data.bytes);
            HIPCHK(hipMemcpyAsync(keep.p, fu.data.p, fu.data.bytes, hipMemcpyDeviceToDevice, c->stream));
            for (int l = 0; l < c->nlevels; l++) {
                LevelArgs<T> a = level_args<T>(c, l, s, op, 1.0);
                if (!a.nblk) continue;
                float best = 1e30f;
                int bestk = 0;
                for (int k = 0; k < o.NCFG; k++) {
                    // reps sweeps captured in a graph and replayed, as a solve
                    // runs them: eager launches of the few-microsecond sweeps of
                    // small levels would time the host's launch rate instead;
                    // the faster of two replays counts
                    o.sweep(fu, field(c, b), a, l, k);
                    Graph gr;
                    capture_graph(c, gr, [&] {
                        for (int i = 0; i < reps; i++) o.sweep(fu, field(c, b), a, l, k);
                    });
                    HIPCHK(hipGraphLaunch(gr.x, c->stream));
                    float ms = 1e30f;
                    for (int trial = 0; trial < 2; trial++) {
                        HIPCHK(hipEventRecord(e0, c->stream));
                        HIPCHK(hipGraphLaunch(gr.x, c->stream));
                        HIPCHK(hipEventRecord(e1, c->stream));
                        HIPCHK(hipEventSynchronize(e1));
                        float t = 0.f;
                        HIPCHK(hipEventElapsedTime(&t, e0, e1));
                        ms = std::min(ms, t);
                    }
                    if (ms < best) {
                        best = ms;
                        bestk = k;
                    }
                }
                c->sweep_cfg[(t == SRMG_F64 ? 0 : c->nlevels) + l] = bestk;
                if (chosen) chosen[l] = bestk;
            }
            std::fill(fu.cur.begin(), fu.cur.end(), 0);
            HIPCHK(hipMemcpyAsync(fu.data.p, keep.p, fu.data.bytes, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        });
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int srmg_set_sweep_config(srmg_ctx *c, int dtype, int level, int cfg) {
    return guarded([&] {
        check_tree(c);
        if (level < 1 || level > c->nlevels) fail("level out of range");
        if (dtype != SRMG_F64 && dtype != SRMG_F32) fail("bad dtype");
        if (cfg < 0 || cfg > 11) fail("sweep configuration must be 0..11");
        c->sweep_cfg[(dtype == SRMG_F64 ? 0 : c->nlevels) + level - 1] = cfg;
    });
}

int srmg_set_group_tiles(srmg_ctx *c, int enable) {
    return guarded([&] {
        if (!c) fail("null context");
        c->group_tiles = enable != 0;
    });
}

int srmg_set_axis_gemm(srmg_ctx *c, int variant) {
    return guarded([&] {
        if (!c) fail("null context");
        if (variant < 0 || variant > 2) fail("axis transform variant must be 0, 1 or 2");
        c->axis_gemm = variant;
    });
}

int srmg_set_group_order(srmg_ctx *c, int order) {
    return guarded([&] {
        if (!c) fail("null context");
        if (order < 0 || order > 1) fail("group order must be 0 or 1");
        c->group_order = order;
    });
}

int srmg_level_grouped(srmg_ctx *c, int level, int *grouped) {
    return guarded([&] {
        check_tree(c);
        if (level < 1 || level > c->nlevels || !grouped) fail("bad arguments");
        *grouped = c->lv[level - 1].grouped ? 1 : 0;
    });
}

int srmg_level_info(srmg_ctx *c, int level, int64_t *nblocks, int64_t *nleaf, int64_t *nslots) {
    return guarded([&] {
        check_tree(c);
        if (level < 1 || level > c->nlevels) fail("level out of range");
        const LevelInfo &L = c->lv[level - 1];
        if (nblocks) *nblocks = L.nblk;
        if (nleaf) *nleaf = L.nleaf;
        if (nslots) *nslots = L.nslot;
    });
}

}  // extern "C"

static void electric_field_impl(srmg_ctx *c, int ui, int eabsi, double h, const double *e0, const int8_t *bc,
                                double *e_host, int G, const srmg_freebc *fb);

extern "C" {

int srmg_electric_field(srmg_ctx *c, int ui, int eabsi, double h, const double *e0, const int8_t *bc,
                        double *e_host, int G) {
    return guarded([&] { electric_field_impl(c, ui, eabsi, h, e0, bc, e_host, G, nullptr); });
}

int srmg_freebc_electric_field(srmg_ctx *c, srmg_freebc *fb, int ui, int eabsi, double h, const double *e0,
                               const int8_t *bc, double *e_host, int G) {
    return guarded([&] {
        if (!fb) fail("null free boundary");
        if (!fb->have_b) fail("srmg_freebc_apply has not been called");
        electric_field_impl(c, ui, eabsi, h, e0, bc, e_host, G, fb);
    });
}

}  // extern "C"

static void electric_field_impl(srmg_ctx *c, int ui, int eabsi, double h, const double *e0, const int8_t *bc,
                                double *e_host, int G, const srmg_freebc *fb) {
    {
        check_tree(c);
        if (!e0 || !bc) fail("null argument");
        if (h <= 0.0) fail("h must be positive");
        if (e_host && G < 1) fail("G must be >= 1");
        FieldBuf &u = field(c, ui), &ea = field(c, eabsi);
        if (u.dtype != SRMG_F64 || ea.dtype != SRMG_F64) fail("u and eabs must be FP64");
        const int D = c->D, M = c->M;
        const int64_t NF = (int64_t)(M + 1) * (D == 2 ? M : M * M);
        DevBuf ef;
        ef.alloc((size_t)c->nb * D * NF * sizeof(double));
        srmg_op op{};
        op.geometry = SRMG_CARTESIAN;
        std::memcpy(op.bc, bc, 27);
        dispatch(c, SRMG_F64, [&](auto &o) {
            using OT = std::remove_reference_t<decltype(o)>;
            using T = std::remove_reference_t<decltype(*o.X(u, 0))>;
            if constexpr (!std::is_same<T, double>::value) return;
            else {
            o.settle(u);
            o.settle(ea);
            for (int l = 0; l < c->nlevels; l++) {
                if (!c->lv[l].nblk) continue;
                LevelArgs<double> a = level_args<double>(c, l, 1.0, &op, 1.0);
                EFieldArgs<double> args;
                args.blk0 = c->lv[l].blk0;
                args.hl = h / (double)(1 << l);
                for (int d = 0; d < 3; d++) args.e0[d] = d < D ? e0[d] : 0.0;
                for (int q = 0; q < 6; q++) args.sg[q] = a.sg[q];
                args.fbb = nullptr;
                args.coord = c->d_coord.as<int>();
                args.fdim = 0;
                args.fkd = 0;
                args.fp = 1;
                if (fb && l + 1 <= fb->level) {
                    args.fbb = fb->b.as<double>();
                    args.fdim = fb->dim;
                    args.fkd = (int)(fb->root[fb->dim] << l);
                    args.fp = 1 << (fb->level - 1 - l);
                }
                OT::launch_efield(c, ea.data.as<double>(), ef.as<double>(), u.data.as<double>(), u.ib.as<double>(),
                                  args, c->lv[l].nblk);
            }
            }
        });
        HIPCHK(hipGetLastError());
        if (e_host) {
            std::vector<double> h_ef((size_t)c->nb * D * NF);
            HIPCHK(hipMemcpyAsync(h_ef.data(), ef.p, h_ef.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            const int64_t Sv = M + 2 * G + 1;
            const int64_t SvD = D == 2 ? Sv * Sv : Sv * Sv * Sv;
            for (int b = 0; b < c->nb; b++) {
                const int64_t hid = c->dev2host[b];  // 1-based host block id
                for (int d = 0; d < D; d++) {
                    const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;
                    for (int64_t i = 0; i < NF; i++) {
                        int cc[3] = {0, 0, 0};
                        cc[d] = (int)(i % (M + 1));
                        cc[da] = (int)((i / (M + 1)) % M);
                        if (D == 3) cc[db] = (int)((i / (M + 1)) / M);
                        int64_t off = 0, st = 1;
                        for (int e = 0; e < D; e++) {
                            off += (int64_t)(cc[e] + G) * st;
                            st *= Sv;
                        }
                        e_host[off + (int64_t)d * SvD + (hid - 1) * D * SvD] = h_ef[((size_t)b * D + d) * NF + i];
                    }
                }
            }
        } else {
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
}

// ---------------------------------------------------------------------------
// exact solves on stacked uniform grids (srmg_grid_solve and its variant
// without device allocations)
// ---------------------------------------------------------------------------

struct GridSolvePlan {
    int dims;      // D
    int N[3];      // extents (N[2] = 1 in 2D)
    size_t total;  // cells per grid
    int rdim;      // radial dimension, or -1
};

// device buffers of one grid solve: per axis the eigenvalues and the two
// transform matrices (each as written by axis_matrix_both), the per-grid
// Helmholtz shifts and two ping-pong grids
struct GridSolveBufs {
    double *lam[3], *P[3], *Q[3];
    double *shift, *g1, *g2;
};

static GridSolvePlan grid_solve_plan(srmg_ctx *c, int D, const int *n, const srmg_op *op, const double *f,
                                     const double *u, int ngrid) {
    if (!c || !n || !op || !f || !u) fail("null argument");
    if (D != 2 && D != 3) fail("D must be 2 or 3");
    if (ngrid < 1) fail("ngrid must be >= 1");
    GridSolvePlan g{D, {n[0], n[1], D == 3 ? n[2] : 1}, 0, -1};
    for (int d = 0; d < D; d++)
        if (g.N[d] < 1) fail("bad grid extents");
    if (op->geometry != SRMG_CARTESIAN && D != 2) fail("cylindrical geometry is two-dimensional");
    g.rdim = (op->geometry == SRMG_CYL1) ? 0 : (op->geometry == SRMG_CYL2 ? 1 : -1);
    g.total = (size_t)g.N[0] * g.N[1] * g.N[2];
    return g;
}

// Separable exact solve of ngrid stacked grids: the operands are computed on
// the host and copied into bufs on the context's stream, then forward
// transforms along every axis, the eigenvalue scaling and the inverse
// transforms, the last one into u. Returns when u is complete.
static void grid_solve_run(srmg_ctx *c, const GridSolvePlan &g, double s, const srmg_op *op, const double *f,
                           double *u, int ngrid, const GridSolveBufs &bufs) {
    const int *N = g.N;
    const size_t total = g.total;
    // host operands stay alive until the stream is synchronized below
    std::vector<std::vector<double>> hold;
    hold.reserve(3 * 3 + 1);
    const int dims = g.dims;
    for (int d = 0; d < dims; d++) {
        int il = 0, ih = 0, m = 1;
        for (int e = 0; e < dims; e++) {
            il += ((e == d ? -1 : 0) + 1) * m;
            ih += ((e == d ? 1 : 0) + 1) * m;
            m *= 3;
        }
        const int nd = N[d];
        std::vector<double> hl(nd), hP((size_t)nd * nd), hQ((size_t)nd * nd);
        axis_decomp(nd, d == g.rdim, (double)op->bc[il], (double)op->bc[ih], hl.data(), hP.data(), hQ.data());
        hold.push_back(std::move(hl));
        HIPCHK(hipMemcpyAsync(bufs.lam[d], hold.back().data(), nd * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        hold.push_back(axis_matrix_both(hP, nd));
        HIPCHK(hipMemcpyAsync(bufs.P[d], hold.back().data(), hold.back().size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
        hold.push_back(axis_matrix_both(hQ, nd));
        HIPCHK(hipMemcpyAsync(bufs.Q[d], hold.back().data(), hold.back().size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    }
    hold.push_back(std::vector<double>(ngrid, op->disc == SRMG_HELMHOLTZ ? -s * op->k2 : 0.0));
    HIPCHK(hipMemcpyAsync(bufs.shift, hold.back().data(), ngrid * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->axis_gemm == 2 && l1_fused_ok(dims, N)) {
        launch_l1_fused<8, double, false>(c->stream, ngrid, N, nullptr, f, nullptr, nullptr, bufs.Q, bufs.P,
                                          bufs.lam, bufs.shift, bufs.g1, bufs.g2, nullptr, u);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        return;
    }
    const double *in = f;
    double *a = bufs.g1, *b2 = bufs.g2;
    auto axis = [&](const double *A, int d, int stride, const double *x, double *y) {
        launch_axis_gemm(c->stream, A, N[d], stride, (int)(ngrid * total / N[d]), x, y, c->axis_gemm);
    };
    int stride = 1;
    for (int d = 0; d < dims; d++) {
        axis(bufs.Q[d], d, stride, in, a);
        in = a;
        std::swap(a, b2);
        stride *= N[d];
    }
    double *cur = const_cast<double *>(in);
    hipLaunchKernelGGL(k_eig_scale, dim3(nwg(total * ngrid)), dim3(WG), 0, c->stream, cur, bufs.lam[0], bufs.lam[1],
                       dims == 3 ? bufs.lam[2] : bufs.lam[1], N[0], N[1], dims, bufs.shift, total, ngrid);
    // inverse transforms in the reverse axis order (the last forward axis first)
    for (int d = dims - 1; d >= 0; d--) {
        stride = 1;
        for (int e = 0; e < d; e++) stride *= N[e];
        double *dst = (d == 0) ? u : a;
        axis(bufs.P[d], d, stride, cur, dst);
        cur = dst;
        std::swap(a, b2);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
}

extern "C" {

int srmg_set_guess(srmg_ctx *c, srmg_unet *net) {
    return guarded([&] {
        if (!c) fail("null context");
        c->guess_net = net;
    });
}

int srmg_guess_fallbacks(srmg_ctx *c, int *n) {
    return guarded([&] {
        if (!c || !n) fail("null argument");
        *n = c->guess_fallbacks;
    });
}

int srmg_grid_solve(srmg_ctx *c, int D, const int *n, double s, const srmg_op *op, const double *f, double *u,
                    int ngrid) {
    return guarded([&] {
        GridSolvePlan g = grid_solve_plan(c, D, n, op, f, u, ngrid);
        DevBuf P[3], Q[3], lam[3], shift, g1, g2;
        GridSolveBufs bufs{};
        for (int d = 0; d < D; d++) {
            lam[d].alloc(g.N[d] * sizeof(double));
            P[d].alloc(2 * (size_t)g.N[d] * g.N[d] * sizeof(double));
            Q[d].alloc(2 * (size_t)g.N[d] * g.N[d] * sizeof(double));
            bufs.lam[d] = lam[d].as<double>();
            bufs.P[d] = P[d].as<double>();
            bufs.Q[d] = Q[d].as<double>();
        }
        shift.alloc(ngrid * sizeof(double));
        g1.alloc(g.total * ngrid * sizeof(double));
        g2.alloc(g.total * ngrid * sizeof(double));
        bufs.shift = shift.as<double>();
        bufs.g1 = g1.as<double>();
        bufs.g2 = g2.as<double>();
        grid_solve_run(c, g, s, op, f, u, ngrid, bufs);
    });
}

int srmg_grid_solve_noalloc(srmg_ctx *c, int D, const int *n, double s, const srmg_op *op, const double *f,
                            double *u, int ngrid) {
    return guarded([&] {
        GridSolvePlan g = grid_solve_plan(c, D, n, op, f, u, ngrid);
        if (!c->scratch) fail("the context has no device scratch address");
        // carve the scratch array: every piece starts on a 256-byte boundary
        char *base = static_cast<char *>(c->scratch);
        size_t off = 0;
        auto take = [&](size_t bytes) {
            double *p = reinterpret_cast<double *>(base + off);
            off += (bytes + 255) & ~(size_t)255;
            return p;
        };
        GridSolveBufs bufs{};
        for (int d = 0; d < D; d++) {
            bufs.lam[d] = take(g.N[d] * sizeof(double));
            bufs.P[d] = take(2 * (size_t)g.N[d] * g.N[d] * sizeof(double));
            bufs.Q[d] = take(2 * (size_t)g.N[d] * g.N[d] * sizeof(double));
        }
        bufs.shift = take(ngrid * sizeof(double));
        bufs.g1 = take(g.total * ngrid * sizeof(double));
        bufs.g2 = take(g.total * ngrid * sizeof(double));
        if (off > NOALLOC_SCRATCH_BYTES)
            fail("srmg_grid_solve_noalloc: the grids need " + std::to_string(off) + " bytes of scratch, above the " +
                 std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte maximum (use srmg_grid_solve)");
        grid_solve_run(c, g, s, op, f, u, ngrid, bufs);
    });
}

int srmg_axis_decomp(int n, int radial, double alow, double ahigh, double *lam, double *P, double *Q) {
    return guarded([&] {
        if (n < 1 || !lam || !P || !Q) fail("bad arguments");
        axis_decomp(n, radial != 0, alow, ahigh, lam, P, Q);
    });
}


int srmg_freebc_create(srmg_ctx *c, srmg_freebc **out, int dim, int level, const int64_t *rootsize, double h,
                       const double *f, int64_t n) {
    return guarded([&] {
        if (!c || !out || !rootsize || !f) fail("null argument");
        *out = nullptr;
        if (c->D != 2) fail("free boundary conditions are two-dimensional (src/freebc.jl:90)");
        if (dim != 1 && dim != 2) fail("dim must be 1 or 2 (the radial dimension of CylindricalGeometry{dim})");
        if (level < 1 || level > 30) fail("bad level");
        if (h <= 0.0) fail("h must be positive");
        auto fb = std::make_unique<srmg_freebc>();
        fb->dim = dim - 1;
        fb->od = 1 - fb->dim;
        fb->level = level;
        fb->root[0] = rootsize[0];
        fb->root[1] = rootsize[1];
        if (fb->root[0] < 1 || fb->root[1] < 1) fail("bad rootsize");
        fb->h = h;
        fb->n = (fb->root[fb->od] << (level - 1)) * c->M;
        if (n != fb->n) fail("f must hold rootsize[other] 2^(level-1) M values");
        if (n > 16384) fail("free boundary longer than 16384 cells");
        auto up = [&](DevBuf &d, const std::vector<double> &v) {
            d.alloc(v.size() * sizeof(double));
            HIPCHK(hipMemcpy(d.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        };
        up(fb->S2T, dst_matrix_t(n, 2));
        up(fb->S3T, dst_matrix_t(n, 3));
        up(fb->f, std::vector<double>(f, f + n));
        fb->u1.alloc((size_t)n * sizeof(double));
        fb->am.alloc((size_t)n * sizeof(double));
        fb->b.alloc((size_t)n * sizeof(double));
        *out = fb.release();
    });
}

int srmg_freebc_destroy(srmg_freebc *fb) {
    return guarded([&] { delete fb; });
}

int srmg_freebc_apply(srmg_ctx *c, srmg_freebc *fb, int qi, int ui, double s) {
    return guarded([&] {
        if (!c || !fb) fail("null argument");
        check_tree(c);
        if (c->D != 2) fail("free boundary conditions are two-dimensional");
        if (c->trees.size() != 1) fail("free boundary conditions need a context holding one tree");
        FieldBuf &q = field(c, qi), &u = field(c, ui);
        if (q.dtype != SRMG_F64 || u.dtype != SRMG_F64) fail("q and u must be FP64");
        const int M = c->M, D = 2, dim = fb->dim, od = fb->od, ND = 2 * D;
        const int64_t n = fb->n;
        if (n != (fb->root[od] << (fb->level - 1)) * M) fail("the free boundary was made for another block size");
        int finest = 0;
        for (int l = 0; l < c->nlevels; l++)
            if (c->lv[l].nblk) finest = l + 1;
        // src/freebc.jl:94-95
        if (finest > fb->level) fail("free boundary at a level below the finest level of the tree");
        if (c->lv[0].ext[dim] != fb->root[dim] || c->lv[0].ext[od] != fb->root[od])
            fail("the free boundary was made for another root domain");
        dispatch(c, SRMG_F64, [&](auto &o) {
            o.settle(u);
            o.settle(q);
        });
        auto cell = [&](int blk, int cd, int co) {  // storage index of local (dim: cd, other: co)
            int x[2];
            x[dim] = cd;
            x[od] = co;
            return (int64_t)blk * c->V + (M == 16 ? Lay<2, 16>::sidx(x[0], x[1], 0) : Lay<2, 8>::sidx(x[0], x[1], 0));
        };
        if (M != 16 && M != 8) fail("unsupported M");
        const int lmax = std::min(fb->level, c->nlevels);
        // getouter!: the finest level's outermost cell for every entry
        std::vector<int64_t> src((size_t)n, -1);
        std::vector<double> hl((size_t)n, 0.0);
        for (int l = 1; l <= lmax; l++) {
            const LevelInfo &L = c->lv[l - 1];
            const int64_t kd = fb->root[dim] << (l - 1);
            const int64_t p = (int64_t)1 << (fb->level - l);
            const double hv = fb->h / (double)(1 << (l - 1));
            for (int bk = L.blk0; bk < L.blk0 + L.nblk; bk++) {
                if (c->h_coord[(size_t)bk * D + dim] != kd) continue;
                const int64_t i0 = (int64_t)(c->h_coord[(size_t)bk * D + od] - 1) * M;
                for (int i = 1; i <= M; i++)
                    for (int64_t j = p * (i0 + i - 1); j < p * (i0 + i); j++) {
                        src[(size_t)j] = cell(bk, M - 1, i - 1);
                        hl[(size_t)j] = hv;
                    }
            }
        }
        for (int64_t j = 0; j < n; j++)
            if (src[(size_t)j] < 0) fail("the outer boundary is not covered at level 1");
        // setouter!: contributions in the reference's order, per target cell
        struct Contrib { int64_t cell; double coef; int64_t bs; int bl; double bd, sl; };
        std::vector<Contrib> cs;
        auto cyl_factor = [](int64_t cdim) { return std::fabs(1.0 + 1.0 / (double)(2 * cdim - 1)); };
        for (int l = 1; l <= lmax; l++) {
            const LevelInfo &L = c->lv[l - 1];
            const double sl = s / (double)(((int64_t)1 << (l - 1)) * ((int64_t)1 << (l - 1)));
            const int64_t p = (int64_t)1 << (fb->level - l);
            const int64_t kd = fb->root[dim] << (l - 1);
            for (int bk = L.blk0; bk < L.blk0 + L.nblk; bk++) {
                if (c->h_nbr[(size_t)bk * ND + 2 * dim + 1] != NB_BOUNDARY) continue;
                const int64_t Bd = c->h_coord[(size_t)bk * D + dim], Bo = c->h_coord[(size_t)bk * D + od];
                const int64_t i0 = (Bo - 1) * M, Jd = (int64_t)M * (Bd - 1) + M;
                for (int i = 1; i <= M; i++)
                    cs.push_back({cell(bk, M - 1, i - 1), 2.0 * cyl_factor(Jd), p * (i0 + i - 1), (int)p,
                                  (double)p, sl});
            }
            for (int k = L.slot0; k < L.slot0 + L.nslot; k++) {
                const int4 sl4 = c->h_slots[(size_t)k];
                const int bk = sl4.x, fd = sl4.y >> 1, side = sl4.y & 1;
                const int64_t Bd = c->h_coord[(size_t)bk * D + dim], Bo = c->h_coord[(size_t)bk * D + od];
                if (Bd != kd) continue;
                const int dir = (fd == od) ? (side ? 1 : -1) : 0;
                const int iv = (dir == -1) ? 1 : M;  // 1-based valid index along the other dimension
                const int64_t j = (Bo - 1) * M + iv;
                const int64_t Jd = (int64_t)M * (Bd - 1) + M, J1d = (Jd - 1) / 2 + 1;
                const int64_t bs = (dir == 1) ? p * j : p * (j - 3) - 1;
                if (bs < 0 || bs + 2 * p > n) fail("free boundary corner outside the boundary vector");
                cs.push_back({cell(bk, M - 1, iv - 1), 2.0 * 0.25 * cyl_factor(J1d), bs, (int)(2 * p),
                              (double)(2 * p), sl});
            }
        }
        std::stable_sort(cs.begin(), cs.end(), [](const Contrib &a, const Contrib &b) { return a.cell < b.cell; });
        std::vector<int64_t> cells, bs;
        std::vector<int> rs, bl;
        std::vector<double> coef, bd, slv;
        for (size_t r = 0; r < cs.size(); r++) {
            if (r == 0 || cs[r].cell != cs[r - 1].cell) {
                cells.push_back(cs[r].cell);
                rs.push_back((int)r);
            }
            coef.push_back(cs[r].coef);
            bs.push_back(cs[r].bs);
            bl.push_back(cs[r].bl);
            bd.push_back(cs[r].bd);
            slv.push_back(cs[r].sl);
        }
        rs.push_back((int)cs.size());
        DevBuf d_src, d_hl, d_cells, d_rs, d_coef, d_bs, d_bl, d_bd, d_sl;
        auto up = [&](DevBuf &d, const void *h, size_t bytes) {
            d.alloc(std::max<size_t>(bytes, 8));
            if (bytes) HIPCHK(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c->stream));
        };
        up(d_src, src.data(), src.size() * sizeof(int64_t));
        up(d_hl, hl.data(), hl.size() * sizeof(double));
        up(d_cells, cells.data(), cells.size() * sizeof(int64_t));
        up(d_rs, rs.data(), rs.size() * sizeof(int));
        up(d_coef, coef.data(), coef.size() * sizeof(double));
        up(d_bs, bs.data(), bs.size() * sizeof(int64_t));
        up(d_bl, bl.data(), bl.size() * sizeof(int));
        up(d_bd, bd.data(), bd.size() * sizeof(double));
        up(d_sl, slv.data(), slv.size() * sizeof(double));
        const unsigned gn = (unsigned)((n + WG - 1) / WG);
        hipLaunchKernelGGL(k_fb_gather, dim3(gn), dim3(WG), 0, c->stream, fb->u1.as<double>(), u.data.as<double>(),
                           d_src.as<int64_t>(), d_hl.as<double>(), n);
        hipLaunchKernelGGL(k_fb_dst, dim3(gn), dim3(WG), 0, c->stream, fb->S2T.as<double>(), fb->u1.as<double>(),
                           fb->am.as<double>(), n, fb->f.as<double>(), 0.0, 1.0);
        hipLaunchKernelGGL(k_fb_dst, dim3(gn), dim3(WG), 0, c->stream, fb->S3T.as<double>(), fb->am.as<double>(),
                           fb->b.as<double>(), n, (const double *)nullptr, fb->h,
                           (double)(1 << (fb->level - 1)));
        if (!cells.empty())
            hipLaunchKernelGGL(k_fb_setouter, dim3((unsigned)((cells.size() + WG - 1) / WG)), dim3(WG), 0,
                               c->stream, q.data.as<double>(), d_cells.as<int64_t>(), d_rs.as<int>(),
                               d_coef.as<double>(), d_bs.as<int64_t>(), d_bl.as<int>(), d_bd.as<double>(),
                               d_sl.as<double>(), fb->b.as<double>(), (int)cells.size());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        fb->have_b = true;
    });
}

int srmg_freebc_boundary(srmg_freebc *fb, double *b, int64_t n) {
    return guarded([&] {
        if (!fb || !b) fail("null argument");
        if (!fb->have_b) fail("srmg_freebc_apply has not been called");
        if (n != fb->n) fail("wrong length");
        HIPCHK(hipMemcpy(b, fb->b.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
