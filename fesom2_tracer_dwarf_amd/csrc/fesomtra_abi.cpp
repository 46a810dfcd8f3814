// C ABI of the tracer advection library (declared in include/fesomtra.h).
//
// Mirrors lib/tracer_c_interface.F90: one process-wide instance, int status codes, copy-in /
// copy-out of Fortran-ordered arrays with 1-based indices.  At tracer_init_complete the mesh is
// validated, renumbered (nodes along a Hilbert curve, elements and edges after their first
// node), transposed to the level-major device layout and uploaded to HBM together with the CSR
// node->edge lists, which keep the reference's edge order so every gathered sum is added in the
// order of the reference's serial scatter.  Tracers live in blocks of kb interleaved tracers
// (tracer id i is lane i % kb of block i / kb); one launch sequence advances the lanes of a block
// that share a scheme.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/fesomtra.h"
#include "fesomtra_restart.hpp"
#include "fesomtra_meshfile.hpp"
#include "fesomtra_kernels.hpp"
#include "fesomtra_tile.hpp"

using namespace fesomtra;

namespace {

struct SchemeSel {
    int hor = HOR_UPW1, ver = VER_UPW1;
    bool fct = false;
    double ph = 1.0, pv = 1.0;
    bool operator==(const SchemeSel &o) const {
        return hor == o.hor && ver == o.ver && fct == o.fct && ph == o.ph && pv == o.pv;
    }
};

struct State {
    int nl = 0, N = 0, E = 0, T = 0, max_adj = 0;
    bool dims = false;
    std::vector<double> coord;
    std::vector<int> enodes, edges, etri, nie, nie_num;
    std::vector<int> nlev_n, ulev_n, nlev_e, ulev_e, nbl, updn;
    std::vector<double> elem_area, area, areasvol, edxdy, ecd, elem_cos, gsca;
    std::vector<double> hnode, hnode_new, helem, Z, zbar3, zbar;
    std::vector<double> uv, w, wi, we;
    // Ensemble: nmem velocities on this mesh; member 0 is uv / w / wi / we, members 1.. are kept
    // here in the same layouts, and each tracer belongs to one member (tr_mem, 0-based)
    int nmem = 1;
    std::vector<std::vector<double>> muv, mw, mwi, mwe;
    std::vector<int> tr_mem;
    std::vector<double> node_dt;  // time-step factor per node (empty: 1 everywhere)
    bool have_coord = false, have_conn = false, have_nlev = false, have_elev = false;
    bool have_geom = false, have_vert = false, have_hnode = false, have_muscl = false;
    bool have_vel = false, have_zbar = false;
    int myDim = -1;
    int ntr = 0;
    std::vector<std::vector<double>> values;
    std::vector<SchemeSel> schemes;
    bool initialized = false;  // the reference's "tracer_initialized"
    int prec = PREC_F64;
    int kb_req = 0;  // requested tracer block width (0: automatic)
    int kb = 1;      // block width of the device storage
    LaunchCfg cfg{};
    bool use_graph = false;
    bool stochastic = true;  // stochastic rounding of the update in the single-precision modes
    bool compensated = false;  // single precision: keep each stored value's rounding remainder
    uint32_t step_seed = 0;  // advanced once per step
    uint32_t ckpt_seed = 0;  // step_seed at the last tracer_checkpoint(0)
    uint64_t gen = 0;  // bumped whenever captured launch sequences become stale

    // device side
    bool on_device = false;
    hipStream_t stream = nullptr;
    std::vector<int> perm_n, inv_n, perm_t, inv_t, perm_e, inv_e;
    size_t Np = 0, Tp = 0, Ep = 0;
    MeshCol<double> m64{};  // mesh and velocities stay FP64 in every precision mode
    // device-order host tables kept for the tile builder
    std::vector<int> h_ptr, h_e2, h_top, h_oth, h_nie, h_nie_cnt, h_en, h_nlev, h_nle;
    std::vector<int2> h_ud;
    std::vector<int4> h_ne;
    // fused single-pass path (FP64, FCT, one tracer per block)
    bool fused = false;
    bool wsplit = false;  // dynamics%use_wsplit
    RestartFiles rs;            // restart files the state was read from (tracer_finalize rewrites them)
    bool from_restart = false;
    int ab_mode = 0;      // valuesAB per step: 0 := values, 1 frozen, 2 AB2 interpolation
    double ab_eps = 0.1;  // AB2 offset epsilon
    std::vector<void *> toldbuf;                    // per block: valuesold (ab_mode 2)
    std::vector<void *> tabbuf2;                    // per block: valuesAB of the next step (ab_mode 2)
    bool ab_fresh = true;  // ab_mode 2: valuesAB to be formed from values and valuesold first
    std::vector<std::vector<double>> ab_host, old_host;  // valuesAB / valuesold across rebuilds
    std::vector<void *> ckpt_ab, ckpt_old;          // per block: copies kept by tracer_checkpoint
    int ckpt_ab_mode = 0;
    bool tiles_built = false;
    // halo tiles of the node kernels (tracer blocks of width <= 4)
    bool halo_tiles = true;
    int ht_built = 0;  // number of node tiles built for the current block width (0: none)
    int et_built = 0;  // number of edge tiles built for the current block width (0: none)
    int rt_built = 0;  // node tiles with two-ring halos (0: none, or they did not fit)
    bool fused_limiter = false;  // limiter bounds and update in one kernel where the tiles allow it
    bool concurrent = true;  // low-order kernels on an aux stream beside the edge-flux kernels
    TileMesh tm{};
    std::vector<void *> dT2;    // second tracer buffer per block (output of a fused step)
    std::vector<char> t2_sync;  // dT2 holds the same below-bottom values as dT
    std::vector<void *> allocs;
    std::vector<void *> dT, tabbuf;  // per tracer block, storage type of the precision mode
    std::vector<void *> dR;          // per block: remainders of the stored values (compensated)
    std::vector<char> tab_on;        // per block: valuesAB held apart from values
    std::vector<void *> ckpt;        // per block: device copy kept by tracer_checkpoint
    std::vector<void *> ckpt_res;    // per block: the remainders kept with it (compensated storage)
    void *adfh = nullptr, *LO = nullptr, *Rp = nullptr, *Rm = nullptr, *adfv = nullptr;
    void *LOe = nullptr;  // compensated single precision: remainder of the stored LO increment
    void *mg = nullptr;   // node-mean gradients of a pass (MeshCol::mg_* entries x 2 x kb)
    double *acc_h = nullptr, *acc_v = nullptr;
    // tracer_zero_tendencies was called and no drop-in call has accumulated since: the next call's
    // update writes its sums instead of adding to the arrays (the cells it does not write, below
    // the bottom and above a cavity's top, were zeroed when the arrays were allocated and are never
    // written)
    bool acc_zero = false;
    void *dbg_h = nullptr, *dbg_v = nullptr;
    bool keep_work = true;  // the drop-in stores the limited fluxes for tracer_get_work
    bool work_kept = false;  // the last drop-in call stored them
    void *tri = nullptr;      // (c', t') of the implicit vertical solve (allocated with wsplit)
    void *cqv = nullptr;      // cavity meshes: the limiter's element-bound rows above element tops
    int cq_rows = 0;          // number of those rows (levels above the slot elements' tops)
    double *stage = nullptr;  // reference-layout staging buffer (N * L doubles)
    int *d_perm_n = nullptr;
    double4 *part = nullptr;
    bool last_fct = false;
    int last_lane = 0;
    uint64_t q_vel = ~0ull;  // velocity version the edge transport m.q was computed from
    uint64_t vel_gen = 0;    // bumped whenever the device velocity changes
    hipGraphExec_t gexec = nullptr;  // captured advect_steps sequence
    int g_nsteps = -1;
    double g_dt = 0.0;
    uint64_t g_gen = 0;
    double last_ms = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool kernel_timing = false;
    double stage_ms[4] = {0, 0, 0, 0};
    long stage_count = 0;
    double qflux_ms = 0.0, step_ms = 0.0;  // summed k_qflux and whole-step times (kernel timing)
    long step_count = 0;
    uint32_t *d_seed = nullptr;  // device copy of step_seed, advanced by a kernel per step
    int *d_lane_mem = nullptr;   // ensembles: member of each block lane, [blocks][kb]
    bool dropin_valid = false;   // the work arrays hold the last tracer_do_oce_adv_tra call's
    bool lo_dirty = false;       // the LO buffer held a values snapshot (single precision, no limiter)
};

State *g = nullptr;
std::string g_err;
constexpr int kStatsBlocks = 1024;

State &S() {
    if (!g) g = new State();
    return *g;
}

int fail(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    fprintf(stderr, "### error: %s\n", buf);
    return -1;
}

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) return fail("%s failed: %s", #call, hipGetErrorString(_e)); \
    } while (0)

size_t storage_bytes(int prec) { return prec == PREC_F64 ? 8 : (prec == PREC_F32 ? 4 : 2); }
size_t compute_bytes(int prec) { return prec == PREC_F64 ? 8 : 4; }

int dalloc_bytes(State &s, void **p, size_t bytes) {
    void *q = nullptr;
    bytes = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    s.allocs.push_back(q);
    *p = q;
    return 0;
}
template <class T>
int dalloc(State &s, T **p, size_t count) {
    void *q;
    if (dalloc_bytes(s, &q, count * sizeof(T))) return -1;
    *p = static_cast<T *>(q);
    return 0;
}
template <class T>
int upload(State &s, const T **p, const std::vector<T> &h) {
    T *q;
    if (dalloc(s, &q, h.size())) return -1;
    if (!h.empty()) HIPCHK(hipMemcpy(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *p = q;
    return 0;
}

void drop_graph(State &s) {
    if (s.gexec) (void)hipGraphExecDestroy(s.gexec);
    s.gexec = nullptr;
    s.g_nsteps = -1;
}

void release_device(State &s) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    drop_graph(s);
    for (void *p : s.allocs) (void)hipFree(p);
    s.allocs.clear();
    s.dT.clear();
    s.dR.clear();
    s.tabbuf.clear();
    s.tab_on.clear();
    s.ckpt.clear();
    s.ckpt_res.clear();
    s.toldbuf.clear();
    s.tabbuf2.clear();
    s.ckpt_ab.clear();
    s.ckpt_old.clear();
    s.dT2.clear();
    s.t2_sync.clear();
    s.d_seed = nullptr;
    s.dropin_valid = false;
    s.q_vel = ~0ull;
    s.tiles_built = false;
    s.tm = TileMesh{};
    s.ht_built = 0;
    s.et_built = 0;
    s.rt_built = 0;
    s.adfh = s.LO = s.Rp = s.Rm = s.adfv = nullptr;
    s.acc_h = s.acc_v = nullptr;
    s.dbg_h = s.dbg_v = nullptr;
    s.tri = nullptr;
    s.stage = nullptr;
    s.d_perm_n = nullptr;
    s.part = nullptr;
    s.on_device = false;
    s.m64 = MeshCol<double>{};

}

int check_dims(const State &s, int nl, int nod2D, const char *who) {
    if (!s.dims) return fail("%s: call tracer_set_mesh_dims first", who);
    if (nl != s.nl) return fail("%s: nl=%d does not match mesh nl=%d", who, nl, s.nl);
    if (nod2D != s.N) return fail("%s: nod2D=%d does not match mesh nod2D=%d", who, nod2D, s.N);
    return 0;
}

bool parse_hor(const char *h, int &v) {
    if (!strcmp(h, "UPW1")) v = HOR_UPW1;
    else if (!strcmp(h, "MUSCL")) v = HOR_MUSCL;
    else if (!strcmp(h, "MFCT")) v = HOR_MFCT;
    else return false;
    return true;
}
bool parse_ver(const char *h, int &v) {
    if (!strcmp(h, "UPW1")) v = VER_UPW1;
    else if (!strcmp(h, "QR4C")) v = VER_QR4C;
    else if (!strcmp(h, "CDIFF")) v = VER_CDIFF;
    else if (!strcmp(h, "PPM")) v = VER_PPM;
    else return false;
    return true;
}

// Hilbert index of (x, y) on a 2^16 grid.
uint64_t hilbert(uint32_t x, uint32_t y) {
    const uint32_t n = 1u << 16;
    uint64_t d = 0;
    for (uint32_t sbit = n / 2; sbit > 0; sbit /= 2) {
        const uint32_t rx = (x & sbit) > 0, ry = (y & sbit) > 0;
        d += (uint64_t)sbit * sbit * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// Node order: Hilbert curve over the coordinates, else breadth-first over the edge graph.
std::vector<int> node_order(const State &s) {
    const int N = s.N;
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    if (s.have_coord) {
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int n = 0; n < N; ++n) {
            x0 = std::min(x0, s.coord[2 * n]);
            x1 = std::max(x1, s.coord[2 * n]);
            y0 = std::min(y0, s.coord[2 * n + 1]);
            y1 = std::max(y1, s.coord[2 * n + 1]);
        }
        const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
        std::vector<uint64_t> key(N);
        for (int n = 0; n < N; ++n)
            key[n] = hilbert((uint32_t)((s.coord[2 * n] - x0) * sx), (uint32_t)((s.coord[2 * n + 1] - y0) * sy));
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
        return order;
    }
    std::vector<int> ptr(N + 1, 0), adj(2 * (size_t)s.E);
    for (int e = 0; e < s.E; ++e) {
        ptr[s.edges[2 * e] + 1]++;
        ptr[s.edges[2 * e + 1] + 1]++;
    }
    for (int n = 0; n < N; ++n) ptr[n + 1] += ptr[n];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < s.E; ++e) {
        adj[fill[s.edges[2 * e]]++] = s.edges[2 * e + 1];
        adj[fill[s.edges[2 * e + 1]]++] = s.edges[2 * e];
    }
    std::vector<char> seen(N, 0);
    order.clear();
    for (int start = 0; start < N; ++start) {
        if (seen[start]) continue;
        seen[start] = 1;
        size_t head = order.size();
        order.push_back(start);
        while (head < order.size()) {
            const int u = order[head++];
            for (int i = ptr[u]; i < ptr[u + 1]; ++i)
                if (!seen[adj[i]]) {
                    seen[adj[i]] = 1;
                    order.push_back(adj[i]);
                }
        }
    }
    return order;
}

int validate(State &s) {
    if (!s.dims) return fail("mesh dimensions not set (tracer_set_mesh_dims)");
    if (!s.have_conn) return fail("connectivity not set (tracer_set_connectivity)");
    if (!s.have_nlev) return fail("node levels not set (tracer_set_levels)");
    if (!s.have_elev) return fail("element levels not set (tracer_set_elem_levels)");
    if (!s.have_geom) return fail("geometry not set (tracer_set_geometry)");
    if (!s.have_hnode) return fail("layer thickness not set (tracer_set_vertical_grid / tracer_set_thickness)");
    if (!s.have_vert) return fail("vertical grid Z_3d_n / zbar_3d_n / helem not set (tracer_set_vertical_grid)");
    if (s.myDim >= 0 && s.myDim != s.N) return fail("myDim_nod2D=%d must equal nod2D=%d", s.myDim, s.N);
    if (s.nl > 255) return fail("nl=%d: at most 255 levels", s.nl);
    // ulevels > 1 (ice-shelf cavities): a column starts at its top layer; it keeps two layers
    for (int n = 0; n < s.N; ++n) {
        if (s.nlev_n[n] < 3 || s.nlev_n[n] > s.nl) return fail("node %d: nlevels_nod2D=%d outside [3, nl]", n + 1, s.nlev_n[n]);
        if (s.ulev_n[n] < 1 || s.nlev_n[n] - s.ulev_n[n] < 2)
            return fail("node %d: ulevels_nod2D=%d leaves fewer than two layers above nlevels_nod2D=%d", n + 1,
                        s.ulev_n[n], s.nlev_n[n]);
    }
    for (int t = 0; t < s.T; ++t) {
        if (s.nlev_e[t] < 2 || s.nlev_e[t] > s.nl) return fail("element %d: nlevels=%d outside [2, nl]", t + 1, s.nlev_e[t]);
        if (s.ulev_e[t] < 1 || s.ulev_e[t] >= s.nlev_e[t])
            return fail("element %d: ulevels=%d leaves no layer above nlevels=%d", t + 1, s.ulev_e[t], s.nlev_e[t]);
        for (int c = 0; c < 3; ++c) {
            const int n = s.enodes[3 * t + c];
            if (s.nlev_n[n] < s.nlev_e[t])
                return fail("node %d has fewer levels (%d) than element %d (%d)", n + 1, s.nlev_n[n], t + 1, s.nlev_e[t]);
            if (s.ulev_n[n] > s.ulev_e[t])
                return fail("node %d starts deeper (ulevels_nod2D=%d) than element %d (ulevels=%d)", n + 1, s.ulev_n[n],
                            t + 1, s.ulev_e[t]);
        }
    }
    for (int e = 0; e < s.E; ++e) {
        const int t1 = s.etri[2 * e], t2 = s.etri[2 * e + 1];
        if (t1 < 0) return fail("edge %d: edge_tri(1) must be a valid element", e + 1);
        // the scatter range of an edge is one interval of levels: its elements' wet layers meet
        if (t2 >= 0 && (s.ulev_e[t2] > s.nlev_e[t1] || s.ulev_e[t1] > s.nlev_e[t2]))
            return fail("edge %d: the wet layers of its elements %d and %d do not meet", e + 1, t1 + 1, t2 + 1);
    }
    return 0;
}

bool has_cavities(const State &s) {
    return std::any_of(s.ulev_e.begin(), s.ulev_e.end(), [](int u) { return u > 1; }) ||
           std::any_of(s.ulev_n.begin(), s.ulev_n.end(), [](int u) { return u > 1; });
}

// nboundary_lay as muscl_adv_init derives it (lib/oce_muscl_adv.F90:96-159).
void derive_nboundary_lay(State &s) {
    s.nbl.assign(s.N, s.nl - 1);
    for (int e = 0; e < s.E; ++e) {
        const int a = s.edges[2 * e], b = s.edges[2 * e + 1];
        const int t1 = s.etri[2 * e], t2 = s.etri[2 * e + 1];
        if (t1 < 0 || t2 < 0) {
            s.nbl[a] = 0;
            s.nbl[b] = 0;
        } else {
            const int lev = std::min(s.nlev_e[t1], s.nlev_e[t2]) - 1;
            s.nbl[a] = std::min(s.nbl[a], lev);
            s.nbl[b] = std::min(s.nbl[b], lev);
        }
    }
}

inline size_t pad64(size_t n) { return (n + 63) / 64 * 64; }

int pow2ceil(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}
// Tracer block width: the requested one, else all tracers in one block (up to kMaxKB).
int block_width(const State &s) { return s.kb_req > 0 ? s.kb_req : std::min(kMaxKB, pow2ceil(std::max(s.ntr, 1))); }
int n_blocks(const State &s) { return (s.ntr + s.kb - 1) / s.kb; }

// Device copy of a reference (rows, ld) array, transposed to [ld][rows_padded] in device order.
template <class M>
std::vector<M> level_major(const std::vector<double> &src, int ld, const std::vector<int> &perm, size_t pad) {
    std::vector<M> h((size_t)ld * pad, M(0));
    for (size_t r = 0; r < perm.size(); ++r) {
        const double *row = &src[(size_t)perm[r] * ld];
        for (int k = 0; k < ld; ++k) h[(size_t)k * pad + r] = (M)row[k];
    }
    return h;
}
// Device copy of a reference (rows, ld) array kept entity-major, rows in device order.
template <class M>
std::vector<M> rows_of(const std::vector<double> &src, int ld, const std::vector<int> &perm) {
    std::vector<M> h((size_t)ld * perm.size());
    for (size_t r = 0; r < perm.size(); ++r)
        for (int k = 0; k < ld; ++k) h[r * ld + k] = (M)src[(size_t)perm[r] * ld + k];
    return h;
}

// Velocities of ensemble member k (0: the shared / first member).
const std::vector<double> &mem_uv(const State &s, int k) { return k == 0 ? s.uv : s.muv[k - 1]; }
const std::vector<double> &mem_w(const State &s, int k) { return k == 0 ? s.w : s.mw[k - 1]; }
const std::vector<double> &mem_wi(const State &s, int k) { return k == 0 ? s.wi : s.mwi[k - 1]; }
const std::vector<double> &mem_we(const State &s, int k) { return k == 0 ? s.we : s.mwe[k - 1]; }

// Device layouts of every member's velocity: uv member-major [nmem][L][Tp][2]; w, w_e, w_i
// member-interleaved [nl][Np][nmem] (one member: the plain level-major layouts).  w_e and w_i
// are left empty where they equal w for every member (the device then reads w for them).
struct VelHost {
    std::vector<double> uv, w, we, wi;
};
VelHost velocity_layout(const State &s) {
    const int L = s.nl - 1, nl = s.nl, M = s.nmem;
    VelHost v;
    v.uv.assign((size_t)M * L * s.Tp * 2, 0.0);
    for (int k = 0; k < M; ++k) {
        const std::vector<double> &src = mem_uv(s, k);
        double *dst = &v.uv[(size_t)k * L * s.Tp * 2];
        for (int t = 0; t < s.T; ++t) {
            const double *row = &src[(size_t)s.perm_t[t] * L * 2];
            for (int z = 0; z < L; ++z) {
                dst[((size_t)z * s.Tp + t) * 2] = row[2 * z];
                dst[((size_t)z * s.Tp + t) * 2 + 1] = row[2 * z + 1];
            }
        }
    }
    auto inter = [&](const std::vector<double> &(*get)(const State &, int)) {
        std::vector<double> out((size_t)nl * s.Np * M, 0.0);
        for (int k = 0; k < M; ++k) {
            const std::vector<double> &src = get(s, k);
            for (int n = 0; n < s.N; ++n) {
                const double *col = &src[(size_t)s.perm_n[n] * nl];
                for (int j = 0; j < nl; ++j) out[((size_t)j * s.Np + n) * M + k] = col[j];
            }
        }
        return out;
    };
    bool we_w = true, wi_w = true;
    for (int k = 0; k < M; ++k) {
        we_w = we_w && mem_we(s, k) == mem_w(s, k);
        wi_w = wi_w && mem_wi(s, k) == mem_w(s, k);
    }
    v.w = inter(mem_w);
    if (!we_w) v.we = inter(mem_we);
    if (!wi_w) v.wi = inter(mem_wi);
    return v;
}

template <class M>
int upload_mesh(State &s, MeshCol<M> &m, const std::vector<int> &ints_n1, const std::vector<int> &ints_n2,
                const std::vector<int> &ints_n3) {
    const int nl = s.nl, L = nl - 1;
    m.N = s.N;
    m.T = s.T;
    m.E = s.E;
    m.nl = nl;
    m.L = L;
    m.Np = (int)s.Np;
    m.Tp = (int)s.Tp;
    m.Ep = (int)s.Ep;
    m.max_adj = s.max_adj;
    if (upload(s, &m.nlev_n, ints_n1) || upload(s, &m.nlev_min, ints_n2) || upload(s, &m.nbl, ints_n3)) return -1;
    // inputs identical bit for bit share one device array (hnode_new = hnode without ALE motion,
    // areasvol = area without cavities), so kernels reading both fetch it once
    auto same_bits = [](const std::vector<double> &x, const std::vector<double> &y) {
        return x.size() == y.size() && (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(double)) == 0);
    };
    if (upload(s, &m.hnode, level_major<M>(s.hnode, L, s.perm_n, s.Np))) return -1;
    if (same_bits(s.hnode_new, s.hnode)) m.hnode_new = m.hnode;
    else if (upload(s, &m.hnode_new, level_major<M>(s.hnode_new, L, s.perm_n, s.Np))) return -1;
    if (upload(s, &m.Z, level_major<M>(s.Z, L, s.perm_n, s.Np))) return -1;
    if (upload(s, &m.area, level_major<M>(s.area, nl, s.perm_n, s.Np))) return -1;
    if (same_bits(s.areasvol, s.area)) m.areasvol = m.area;
    else if (upload(s, &m.areasvol, level_major<M>(s.areasvol, nl, s.perm_n, s.Np))) return -1;
    if (upload(s, &m.zbar, level_major<M>(s.zbar3, nl, s.perm_n, s.Np))) return -1;
    const VelHost vh = velocity_layout(s);
    m.nmem = s.nmem;
    if (upload(s, &m.w, std::vector<M>(vh.w.begin(), vh.w.end()))) return -1;
    m.we = m.w;
    if (!vh.we.empty() && upload(s, &m.we, std::vector<M>(vh.we.begin(), vh.we.end()))) return -1;
    m.wi = m.w;
    if (!vh.wi.empty() && upload(s, &m.wi, std::vector<M>(vh.wi.begin(), vh.wi.end()))) return -1;
    // elements
    std::vector<double> ea(s.T), ec(s.T);
    for (int t = 0; t < s.T; ++t) {
        ea[t] = s.elem_area[s.perm_t[t]];
        ec[t] = s.elem_cos[s.perm_t[t]];
    }
    if (upload(s, &m.elem_area, std::vector<M>(ea.begin(), ea.end()))) return -1;
    if (upload(s, &m.elem_cos, std::vector<M>(ec.begin(), ec.end()))) return -1;
    if (upload(s, &m.gsca, rows_of<M>(s.gsca, 6, s.perm_t))) return -1;
    if (upload(s, &m.helem, level_major<M>(s.helem, L, s.perm_t, s.Tp))) return -1;
    if (upload(s, &m.uv, std::vector<M>(vh.uv.begin(), vh.uv.end()))) return -1;
    if (upload(s, &m.ecd, rows_of<M>(s.ecd, 4, s.perm_e))) return -1;
    if (upload(s, &m.edxdy, rows_of<M>(s.edxdy, 2, s.perm_e))) return -1;
    return 0;
}

// Halo tiles of the node kernels for a tracer block of width kb <= 4: tile t owns the
// tc = 256 / kb consecutive device nodes from t * tc (a compact patch of the space-filling
// curve); its halo is the ascending list of far ends of its nodes' edges outside the tile.  A
// tile's workgroup stages two slots per thread (its column and one halo node), so the halo may
// hold at most tc nodes; meshes with a larger halo anywhere run the untiled kernels.
// Tiles of the fused limiter kernel: tile t owns the tcs = 128 / kb consecutive device nodes
// from t * tcs; its first ring (the far ends of its nodes' edges outside it, sorted) takes the
// node threads after the tile's (at most tcs of them, 256 / kb threads in all) and its second ring
// (the far ends of the first ring's edges outside the tile and the first ring, sorted) the second
// staging slots (at most 256 / kb).  Per node thread its CSR entries' far nodes as staging slots.
// Tiles whose rings do not fit leave the fused kernel off (rt_built = 0).
int build_limiter_tiles(State &s) {
    MeshCol<double> &m = s.m64;
    m.rt_n = 0;
    s.rt_built = 0;
    const int nt_thr = 256 / s.kb, tcs = 128 / s.kb, N = s.N, nt = (N + tcs - 1) / tcs;
    std::vector<int> rptr(nt + 1, 0), rhalo, nr1v(nt, 0), eoff((size_t)nt * nt_thr, 0), rloc, rlmax(nt, 0),
        rsmax(nt, 0), r1, r2;
    for (int t = 0; t < nt; ++t) {
        const int a0 = t * tcs, a1 = std::min(N, a0 + tcs);
        auto in_tile = [&](int o) { return o >= a0 && o < a1; };
        r1.clear();
        for (int n = a0; n < a1; ++n)
            for (int i = s.h_ptr[n]; i < s.h_ptr[n + 1]; ++i)
                if (!in_tile(s.h_oth[i])) r1.push_back(s.h_oth[i]);
        std::sort(r1.begin(), r1.end());
        r1.erase(std::unique(r1.begin(), r1.end()), r1.end());
        if ((int)r1.size() > nt_thr - tcs) return 0;  // does not fit: the unfused kernels run
        auto in_r1 = [&](int o) { return std::binary_search(r1.begin(), r1.end(), o); };
        r2.clear();
        for (int h : r1)
            for (int i = s.h_ptr[h]; i < s.h_ptr[h + 1]; ++i) {
                const int o = s.h_oth[i];
                if (!in_tile(o) && !in_r1(o)) r2.push_back(o);
            }
        std::sort(r2.begin(), r2.end());
        r2.erase(std::unique(r2.begin(), r2.end()), r2.end());
        if ((int)r2.size() > nt_thr) return 0;
        auto slot = [&](int o) {
            if (in_tile(o)) return o - a0;
            if (in_r1(o)) return tcs + (int)(std::lower_bound(r1.begin(), r1.end(), o) - r1.begin());
            return nt_thr + (int)(std::lower_bound(r2.begin(), r2.end(), o) - r2.begin());
        };
        int lt = 0, ls = 0;
        auto entries = [&](int i, int n) {
            eoff[(size_t)t * nt_thr + i] = (int)rloc.size();
            for (int k = s.h_ptr[n]; k < s.h_ptr[n + 1]; ++k) rloc.push_back(slot(s.h_oth[k]));
        };
        for (int n = a0; n < a1; ++n) {
            lt = std::max(lt, s.h_nlev[n] - 1);
            entries(n - a0, n);
        }
        ls = lt;
        for (int k = 0; k < (int)r1.size(); ++k) {
            ls = std::max(ls, s.h_nlev[r1[k]] - 1);
            entries(tcs + k, r1[k]);
        }
        rlmax[t] = lt;
        rsmax[t] = ls;
        nr1v[t] = (int)r1.size();
        rhalo.insert(rhalo.end(), r1.begin(), r1.end());
        rhalo.insert(rhalo.end(), r2.begin(), r2.end());
        rptr[t + 1] = (int)rhalo.size();
    }
    if (rhalo.empty()) rhalo.push_back(0);
    if (rloc.empty()) rloc.push_back(0);
    if (upload(s, &m.rt_hptr, rptr) || upload(s, &m.rt_halo, rhalo) || upload(s, &m.rt_nr1, nr1v) ||
        upload(s, &m.rt_eoff, eoff) || upload(s, &m.rt_loc, rloc) || upload(s, &m.rt_lmax, rlmax) ||
        upload(s, &m.rt_smax, rsmax))
        return -1;
    s.rt_built = nt;
    m.rt_n = (s.halo_tiles && s.fused_limiter) ? nt : 0;
    return 0;
}

int build_halo_tiles(State &s) {
    MeshCol<double> &m = s.m64;
    m.ht_n = m.ht_kb = 0;
    m.rt_n = 0;
    s.ht_built = 0;
    s.rt_built = 0;
    if (s.kb > 4 || s.N < 1) return 0;
    const int tc = 256 / s.kb, N = s.N, nt = (N + tc - 1) / tc;
    std::vector<int> hptr(nt + 1, 0), halo, loc(s.h_oth.size()), lmax(nt, 0), hl;
    for (int t = 0; t < nt; ++t) {
        const int a0 = t * tc, a1 = std::min(N, a0 + tc);
        hl.clear();
        for (int n = a0; n < a1; ++n) {
            lmax[t] = std::max(lmax[t], s.h_nlev[n] - 1);
            for (int i = s.h_ptr[n]; i < s.h_ptr[n + 1]; ++i)
                if (s.h_oth[i] < a0 || s.h_oth[i] >= a1) hl.push_back(s.h_oth[i]);
        }
        std::sort(hl.begin(), hl.end());
        hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
        if ((int)hl.size() > tc) return 0;
        for (int n = a0; n < a1; ++n)
            for (int i = s.h_ptr[n]; i < s.h_ptr[n + 1]; ++i) {
                const int o = s.h_oth[i];
                loc[i] = (o >= a0 && o < a1) ? o - a0 : tc + (int)(std::lower_bound(hl.begin(), hl.end(), o) - hl.begin());
            }
        halo.insert(halo.end(), hl.begin(), hl.end());
        hptr[t + 1] = (int)halo.size();
    }
    if (halo.empty()) halo.push_back(0);  // keep a valid allocation
    if (upload(s, &m.ht_loc, loc) || upload(s, &m.ht_hptr, hptr) || upload(s, &m.ht_halo, halo) ||
        upload(s, &m.ht_lmax, lmax))
        return -1;
    // the edges of each tile's nodes (sorted), whose transports the low-order kernel stages per
    // level (at most 1024 per tile at width 1, 512 at widths 2 and 4; a mesh where a tile has more
    // runs untiled), and per CSR entry its edge's slot in that list (* 2 + whether the node is the
    // edge's second end)
    {
        const int qs = s.kb == 1 ? 4 : 2;
        std::vector<int> eptr(nt + 1, 0), edges, eloc(s.h_e2.size()), el;
        bool fits = true;
        for (int t = 0; t < nt && fits; ++t) {
            const int a0 = t * tc, a1 = std::min(N, a0 + tc);
            el.clear();
            for (int i = s.h_ptr[a0]; i < s.h_ptr[a1]; ++i) el.push_back(s.h_e2[i] >> 1);
            std::sort(el.begin(), el.end());
            el.erase(std::unique(el.begin(), el.end()), el.end());
            if ((int)el.size() > qs * 256) fits = false;
            for (int i = s.h_ptr[a0]; i < s.h_ptr[a1]; ++i)
                eloc[i] = ((int)(std::lower_bound(el.begin(), el.end(), s.h_e2[i] >> 1) - el.begin()) << 1) |
                          (s.h_e2[i] & 1);
            edges.insert(edges.end(), el.begin(), el.end());
            eptr[t + 1] = (int)edges.size();
        }
        m.ht_eptr = m.ht_edges = m.ht_eloc = nullptr;
        if (!fits) return 0;
        if (edges.empty()) edges.push_back(0);
        if (upload(s, &m.ht_eptr, eptr) || upload(s, &m.ht_edges, edges) || upload(s, &m.ht_eloc, eloc)) return -1;
    }
    s.ht_built = nt;
    m.ht_kb = s.kb;
    m.ht_n = s.halo_tiles ? nt : 0;
    return build_limiter_tiles(s);
}

// Edge tiles of the edge-flux kernel for a tracer block of width kb <= 4: tile t owns the
// te = 256 / kb consecutive device edges from t * te (edges are ordered by their smaller node,
// so a tile is a compact patch); its node list holds every node the tile's edges read at the
// levels where the up- and down-wind triangles are both wet (the two ends and the triangles'
// corners), at most 2 * te nodes (two staging slots per thread), else the edge kernel runs
// untiled.  Per edge the eight slots are packed two to a word.
int build_edge_tiles(State &s) {
    MeshCol<double> &m = s.m64;
    m.et_n = m.et_kb = 0;
    s.et_built = 0;
    if (s.kb > 4 || s.E < 1) return 0;
    const int te = 256 / s.kb, E = s.E, nt = (E + te - 1) / te;
    std::vector<int> nptr(nt + 1, 0), nodes, loc((size_t)4 * E, 0), lmax(nt, 0), nl;
    for (int t = 0; t < nt; ++t) {
        const int a0 = t * te, a1 = std::min(E, a0 + te);
        nl.clear();
        for (int e = a0; e < a1; ++e) {
            const int4 q = s.h_ne[e];
            lmax[t] = std::max(lmax[t], std::max(s.h_nle[q.z] - 1, q.w >= 0 ? s.h_nle[q.w] - 1 : 0));
            nl.push_back(q.x);
            nl.push_back(q.y);
            const int2 ud = s.h_ud[e];
            if (ud.x >= 0 && ud.y >= 0)
                for (int c = 0; c < 3; ++c) {
                    nl.push_back(s.h_en[3 * (size_t)ud.x + c]);
                    nl.push_back(s.h_en[3 * (size_t)ud.y + c]);
                }
        }
        std::sort(nl.begin(), nl.end());
        nl.erase(std::unique(nl.begin(), nl.end()), nl.end());
        if ((int)nl.size() > 2 * te) return 0;
        auto slot = [&](int node) { return (int)(std::lower_bound(nl.begin(), nl.end(), node) - nl.begin()); };
        for (int e = a0; e < a1; ++e) {
            const int4 q = s.h_ne[e];
            const int2 ud = s.h_ud[e];
            int sl[8] = {slot(q.x), slot(q.y), 0, 0, 0, 0, 0, 0};
            if (ud.x >= 0 && ud.y >= 0)
                for (int c = 0; c < 3; ++c) {
                    sl[2 + c] = slot(s.h_en[3 * (size_t)ud.x + c]);
                    sl[5 + c] = slot(s.h_en[3 * (size_t)ud.y + c]);
                }
            for (int i = 0; i < 4; ++i) loc[4 * (size_t)e + i] = sl[2 * i] | (sl[2 * i + 1] << 16);
        }
        nodes.insert(nodes.end(), nl.begin(), nl.end());
        nptr[t + 1] = (int)nodes.size();
    }
    if (upload(s, &m.et_loc, loc) || upload(s, &m.et_nptr, nptr) || upload(s, &m.et_nodes, nodes) ||
        upload(s, &m.et_lmax, lmax))
        return -1;
    s.et_built = nt;
    m.et_kb = s.kb;
    m.et_n = s.halo_tiles ? nt : 0;
    return 0;
}

int pull_values(State &s);

// A setter changed data the device copy is built from: the next use rebuilds it, with the
// tracer values as last advanced on the device.
int invalidate(State &s) {
    if (s.on_device && pull_values(s)) return -1;
    s.on_device = false;
    return 0;
}

// New velocities while the device copy stands: uv, w, w_e and w_i are written into the device
// arrays in place (w_e / w_i get arrays of their own when they first differ from w), so a caller
// that changes the velocity every step keeps its tracers, tiles and captured graph.
int refresh_velocity(State &s) {
    s.vel_gen++;
    if (!s.on_device) return 0;
    MeshCol<double> &m = s.m64;
    const VelHost vh = velocity_layout(s);
    HIPCHK(hipStreamSynchronize(s.stream));
    HIPCHK(hipMemcpy(const_cast<double *>(m.uv), vh.uv.data(), vh.uv.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(const_cast<double *>(m.w), vh.w.data(), vh.w.size() * sizeof(double), hipMemcpyHostToDevice));
    auto part = [&](const std::vector<double> &d, const double *&dst) -> int {
        if (d.empty()) {  // equal to w for every member
            if (dst != m.w) s.gen++;  // kernels and graphs take the pointer
            dst = m.w;
            return 0;
        }
        if (dst == m.w) {
            double *q;
            if (dalloc(s, &q, d.size())) return -1;
            dst = q;
            s.gen++;
        }
        HIPCHK(hipMemcpy(const_cast<double *>(dst), d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    };
    if (part(vh.we, m.we) || part(vh.wi, m.wi)) return -1;
    return 0;
}

// Ensembles: the member of every lane of every tracer block (unused lanes: member 0).
int upload_lane_members(State &s) {
    if (!s.d_lane_mem) return 0;
    const int nb = n_blocks(s);
    std::vector<int> h((size_t)nb * s.kb, 0);
    for (int t = 0; t < s.ntr && t < (int)h.size(); ++t) h[t] = t < (int)s.tr_mem.size() ? s.tr_mem[t] : 0;
    HIPCHK(hipStreamSynchronize(s.stream));
    HIPCHK(hipMemcpy(s.d_lane_mem, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

// Ice-shelf cavities: the level tops in device order and the tables of the limiter's element
// bounds above an element's top (MeshCol::cq_*); all null without cavities.
int upload_cavities(State &s) {
    MeshCol<double> &m = s.m64;
    m.ulev_n = m.ulev_max = m.ulev_e = nullptr;
    m.cq_n = 0;
    m.cq_edge = m.cq_top = m.cq_off = m.cq_ptr = m.cq_ent = nullptr;
    s.cq_rows = 0;
    if (!has_cavities(s)) return 0;
    const int N = s.N, T = s.T, A = s.max_adj;
    std::vector<int> un(N), umax(N), ue(T);
    for (int t = 0; t < T; ++t) ue[t] = s.ulev_e[s.perm_t[t]];
    for (int n = 0; n < N; ++n) {
        const int o = s.perm_n[n];
        un[n] = s.ulev_n[o];
        int mx = 0;
        for (int k = 0; k < s.nie_num[o]; ++k) mx = std::max(mx, s.ulev_e[s.nie[(size_t)o * A + k]]);
        umax[n] = mx;  // ulevels_nod2D_max (lib/oce_mesh.F90:1522-1560)
    }
    if (upload(s, &m.ulev_n, un) || upload(s, &m.ulev_max, umax) || upload(s, &m.ulev_e, ue)) return -1;
    // slots: elements whose top lies below a top of one of their nodes' columns
    std::vector<int> slot_of(T, -1), edge, top, off, ptr(N + 1, 0), ent;
    for (int n = 0; n < N; ++n) {
        const int o = s.perm_n[n];
        for (int k = 0; k < s.nie_num[o]; ++k) {
            const int eo = s.nie[(size_t)o * A + k];  // reference element (0-based)
            if (s.ulev_e[eo] <= s.ulev_n[o]) continue;
            const int t = s.inv_t[eo];
            if (slot_of[t] < 0) {
                // the work array is indexed by the element number as an edge number
                if (eo >= s.E)
                    return fail("element %d: the limiter's element bounds above its top read edge %d of %d", eo + 1,
                                eo + 1, s.E);
                slot_of[t] = (int)edge.size();
                edge.push_back(s.inv_e[eo]);
                top.push_back(s.ulev_e[eo]);
                off.push_back(s.cq_rows);
                s.cq_rows += s.ulev_e[eo] - 1;
            }
            ent.push_back(slot_of[t]);
        }
        ptr[n + 1] = (int)ent.size();
    }
    m.cq_n = (int)edge.size();
    if (m.cq_n == 0) return 0;
    if (upload(s, &m.cq_edge, edge) || upload(s, &m.cq_top, top) || upload(s, &m.cq_off, off) ||
        upload(s, &m.cq_ptr, ptr) || upload(s, &m.cq_ent, ent))
        return -1;
    return 0;
}

// Node-mean gradient entries (MeshCol::mg_*): for every edge with up/down triangles, the levels
// of its range [element top, deepest element bottom] outside the triangles' shared wet range
// [ulevels_nod2D_max, nlevels_nod2D_min - 1] of both ends (all its levels without both triangles)
// take the mean gradients of both end nodes; each node gets one contiguous range of entries
// covering the levels its edges ask for.  nmin: nlevels_nod2D_min in device order.
int upload_mean_grad_levels(State &s, const std::vector<int> &nmin) {
    MeshCol<double> &m = s.m64;
    m.mg_n = 0;
    m.mg_node = m.mg_lev = m.mg_base = nullptr;
    const int N = s.N, E = s.E, A = s.max_adj;
    const bool cav = has_cavities(s);
    std::vector<int> umax(N, 1);
    if (cav)
        for (int n = 0; n < N; ++n) {
            const int o = s.perm_n[n];
            int mx = 0;
            for (int k = 0; k < s.nie_num[o]; ++k) mx = std::max(mx, s.ulev_e[s.nie[(size_t)o * A + k]]);
            umax[n] = mx;
        }
    auto etop = [&](int t) { return cav ? s.ulev_e[s.perm_t[t]] : 1; };
    std::vector<int> lo(N, 1 << 30), hi(N, 0);
    auto want = [&](int n, int a, int b) {
        if (a > b) return;
        lo[n] = std::min(lo[n], a);
        hi[n] = std::max(hi[n], b);
    };
    for (int e = 0; e < E; ++e) {
        const int4 ne = s.h_ne[e];
        const int2 ud = s.h_ud[e];
        const int nl1 = s.h_nle[ne.z] - 1, nl2 = ne.w >= 0 ? s.h_nle[ne.w] - 1 : 0;
        const int top = std::max(nl1, nl2);
        const int et = ne.w >= 0 ? std::min(etop(ne.z), etop(ne.w)) : etop(ne.z);
        int sb = 1, se = 0;
        if (ud.x >= 0 && ud.y >= 0) {
            se = std::min(nmin[ne.x], nmin[ne.y]) - 1;
            if (cav) sb = std::max(umax[ne.x], umax[ne.y]);
        }
        for (int n : {ne.x, ne.y}) {
            if (se < sb) {  // no shared range: every level
                want(n, et, top);
            } else {
                want(n, et, std::min(top, sb - 1));
                want(n, std::max(et, se + 1), top);
            }
        }
    }
    std::vector<int> node, lev, base(s.Np, 0);
    for (int n = 0; n < N; ++n) {
        if (lo[n] > hi[n]) continue;
        base[n] = (int)node.size() - lo[n];
        for (int z = lo[n]; z <= hi[n]; ++z) {
            node.push_back(n);
            lev.push_back(z);
        }
    }
    if (node.size() > (size_t)INT32_MAX / 2) return fail("too many node-mean gradient entries (%zu)", node.size());
    m.mg_n = (int)node.size();
    if (m.mg_n == 0) return 0;
    if (upload(s, &m.mg_node, node) || upload(s, &m.mg_lev, lev) || upload(s, &m.mg_base, base)) return -1;
    return 0;
}

// The per-node time-step factors in device order (m.dtn = null without them).
int upload_node_dt(State &s) {
    s.m64.dtn = nullptr;
    if (s.node_dt.empty()) return 0;
    std::vector<double> h(s.Np, 1.0);
    for (int n = 0; n < s.N; ++n) h[n] = s.node_dt[s.perm_n[n]];
    return upload(s, &s.m64.dtn, h);
}

int build_device(State &s) {
    if (validate(s)) return -1;
    if (!s.have_vel) {  // no velocity given yet: at rest, as the reference's zero-initialised arrays
        s.uv.assign((size_t)2 * (s.nl - 1) * s.T, 0.0);
        s.w.assign((size_t)s.nl * s.N, 0.0);
        s.wi = s.w;
        s.we = s.w;
        s.have_vel = true;
    }
    for (const auto &sc : s.schemes)
        if ((sc.hor == HOR_MUSCL || sc.hor == HOR_MFCT) && !s.have_muscl)
            return fail("MUSCL/MFCT need edge_up_dn_tri (tracer_set_muscl_support)");
    if (!s.have_muscl) derive_nboundary_lay(s);
    release_device(s);
    if (!s.stream) HIPCHK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    if (!s.ev0) HIPCHK(hipEventCreate(&s.ev0));
    if (!s.ev1) HIPCHK(hipEventCreate(&s.ev1));
    // the aux stream of the low-order kernels (default priority, like the main stream)
    if (s.concurrent && !s.cfg.aux) {
        HIPCHK(hipStreamCreateWithFlags(&s.cfg.aux, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&s.cfg.fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&s.cfg.join, hipEventDisableTiming));
    }
    const int N = s.N, T = s.T, E = s.E, nl = s.nl, L = nl - 1, A = s.max_adj;
    s.Np = pad64(N);
    s.Tp = pad64(T);
    s.Ep = pad64(E);

    // Renumbering: nodes along the curve, elements / edges by their smallest new node.
    s.perm_n = node_order(s);
    s.inv_n.assign(N, 0);
    for (int i = 0; i < N; ++i) s.inv_n[s.perm_n[i]] = i;
    {
        std::vector<int64_t> key(T);
        for (int t = 0; t < T; ++t) {
            const int a = s.inv_n[s.enodes[3 * t]], b = s.inv_n[s.enodes[3 * t + 1]], c = s.inv_n[s.enodes[3 * t + 2]];
            key[t] = std::min(a, std::min(b, c));
        }
        s.perm_t.resize(T);
        std::iota(s.perm_t.begin(), s.perm_t.end(), 0);
        std::stable_sort(s.perm_t.begin(), s.perm_t.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    s.inv_t.assign(T, 0);
    for (int i = 0; i < T; ++i) s.inv_t[s.perm_t[i]] = i;
    {
        std::vector<int64_t> key(E);
        for (int e = 0; e < E; ++e) {
            const int a = s.inv_n[s.edges[2 * e]], b = s.inv_n[s.edges[2 * e + 1]];
            key[e] = (int64_t)std::min(a, b) * (int64_t)N + std::max(a, b);
        }
        s.perm_e.resize(E);
        std::iota(s.perm_e.begin(), s.perm_e.end(), 0);
        std::stable_sort(s.perm_e.begin(), s.perm_e.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    s.inv_e.assign(E, 0);
    for (int i = 0; i < E; ++i) s.inv_e[s.perm_e[i]] = i;

    // node integers (lib/oce_mesh.F90:1522-1560 for nlevels_nod2D_min)
    std::vector<int> i_nlev(N), i_nmin(N), i_nbl(N);
    for (int n = 0; n < N; ++n) {
        const int o = s.perm_n[n];
        int mn = 1 << 30;
        for (int k = 0; k < s.nie_num[o]; ++k) mn = std::min(mn, s.nlev_e[s.nie[(size_t)o * A + k]]);
        i_nlev[n] = s.nlev_n[o];
        i_nmin[n] = mn;
        i_nbl[n] = s.nbl[o];
    }
    MeshCol<double> &m = s.m64;
    // connectivity (shared by both precisions)
    {
        std::vector<int> h((size_t)N * A, 0), cnt(N);
        for (int n = 0; n < N; ++n) {
            const int o = s.perm_n[n];
            cnt[n] = s.nie_num[o];
            for (int k = 0; k < cnt[n]; ++k) h[(size_t)n * A + k] = s.inv_t[s.nie[(size_t)o * A + k]];
        }
        if (upload(s, &m.nie, h) || upload(s, &m.nie_num, cnt)) return -1;
        s.h_nie = h;
        s.h_nie_cnt = cnt;
    }
    {
        std::vector<int> en((size_t)3 * T), nle(T);
        for (int t = 0; t < T; ++t) {
            const int o = s.perm_t[t];
            for (int c = 0; c < 3; ++c) en[3 * t + c] = s.inv_n[s.enodes[3 * o + c]];
            nle[t] = s.nlev_e[o];
        }
        if (upload(s, &m.enodes, en) || upload(s, &m.nlev_e, nle)) return -1;
        s.h_en = en;
        s.h_nle = nle;
    }
    std::vector<int> top_of(E), utop_of(E);
    {
        std::vector<int4> ne(E);
        std::vector<int2> ud(E);
        for (int e = 0; e < E; ++e) {
            const int o = s.perm_e[e];
            const int t1 = s.etri[2 * o], t2 = s.etri[2 * o + 1];
            ne[e] = make_int4(s.inv_n[s.edges[2 * o]], s.inv_n[s.edges[2 * o + 1]], s.inv_t[t1], t2 >= 0 ? s.inv_t[t2] : -1);
            if (s.have_muscl) {
                const int u = s.updn[2 * o], d = s.updn[2 * o + 1];
                ud[e] = make_int2(u >= 0 ? s.inv_t[u] : -1, d >= 0 ? s.inv_t[d] : -1);
            } else {
                ud[e] = make_int2(-1, -1);
            }
            // deepest and shallowest level of the scatters' range (lib/oce_adv_tra_driver.F90:125-136)
            top_of[o] = std::max(s.nlev_e[t1] - 1, t2 >= 0 ? s.nlev_e[t2] - 1 : 0);
            utop_of[o] = t2 >= 0 ? std::min(s.ulev_e[t1], s.ulev_e[t2]) : s.ulev_e[t1];
        }
        if (upload(s, &m.edge_ne, ne) || upload(s, &m.updn, ud)) return -1;
        s.h_ud = ud;
        s.h_ne = ne;
    }
    // CSR node -> incident edges in ascending reference edge number (the serial scatter order)
    {
        std::vector<int> ptr(N + 1, 0);
        for (int o = 0; o < E; ++o) {
            ptr[s.inv_n[s.edges[2 * o]] + 1]++;
            ptr[s.inv_n[s.edges[2 * o + 1]] + 1]++;
        }
        for (int n = 0; n < N; ++n) ptr[n + 1] += ptr[n];
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        std::vector<int> e2(ptr[N]), top(ptr[N]), utop(ptr[N]), oth(ptr[N]);
        for (int o = 0; o < E; ++o) {
            const int a = s.inv_n[s.edges[2 * o]], b = s.inv_n[s.edges[2 * o + 1]];
            int i = fill[a]++;
            e2[i] = s.inv_e[o] * 2 + 0;
            top[i] = top_of[o];
            utop[i] = utop_of[o];
            oth[i] = b;
            i = fill[b]++;
            e2[i] = s.inv_e[o] * 2 + 1;
            top[i] = top_of[o];
            utop[i] = utop_of[o];
            oth[i] = a;
        }
        if (upload(s, &m.ne_ptr, ptr) || upload(s, &m.ne_e2, e2) || upload(s, &m.ne_top, top) ||
            upload(s, &m.ne_other, oth))
            return -1;
        m.ne_utop = nullptr;
        if (has_cavities(s) && upload(s, &m.ne_utop, utop)) return -1;
        s.h_ptr = ptr;
        s.h_e2 = e2;
        s.h_top = top;
        s.h_oth = oth;
    }
    s.h_nlev = i_nlev;
    if (upload_mesh<double>(s, s.m64, i_nlev, i_nmin, i_nbl)) return -1;
    if (upload_node_dt(s)) return -1;
    if (upload_cavities(s)) return -1;
    if (upload_mean_grad_levels(s, i_nmin)) return -1;
    if (dalloc(s, &s.m64.q, (size_t)L * s.Ep * s.nmem)) return -1;
    {
        const int *p;
        if (upload(s, &p, s.perm_n)) return -1;
        s.d_perm_n = const_cast<int *>(p);
    }
    // work arrays, interleaved by the block width
    s.kb = block_width(s);
    const size_t cb = compute_bytes(s.prec), sb = storage_bytes(s.prec), kb = s.kb;
    const size_t NLn = (size_t)L * s.Np, ELn = (size_t)L * s.Ep, NIn = (size_t)nl * s.Np;
    // the upwind edge fluxes are recomputed where they are summed, never stored
    if (dalloc_bytes(s, &s.adfh, ELn * kb * cb) ||
        dalloc_bytes(s, &s.LO, NLn * kb * cb) || dalloc_bytes(s, &s.Rp, NLn * kb * cb) ||
        dalloc_bytes(s, &s.Rm, NLn * kb * cb) || dalloc_bytes(s, &s.adfv, NIn * kb * cb))
        return -1;
    if (dalloc(s, &s.part, (size_t)kStatsBlocks) || dalloc(s, &s.stage, (size_t)N * L)) return -1;
    s.mg = nullptr;
    if (s.m64.mg_n > 0 && dalloc_bytes(s, &s.mg, (size_t)s.m64.mg_n * 2 * kb * cb)) return -1;
    if (dalloc(s, &s.d_seed, 64)) return -1;
    s.d_lane_mem = nullptr;
    if (s.nmem > 1) {
        if (dalloc(s, &s.d_lane_mem, (size_t)n_blocks(s) * s.kb)) return -1;
        if (upload_lane_members(s)) return -1;
    }
    s.cqv = nullptr;
    if (s.m64.cq_n > 0) {
        const size_t bytes = (size_t)s.cq_rows * 2 * kb * cb;
        if (dalloc_bytes(s, &s.cqv, bytes)) return -1;
        HIPCHK(hipMemset(s.cqv, 0, std::max<size_t>(bytes, 256)));
    }
    // work arrays start at zero: the kernels write only the levels a mesh entity has, so the
    // levels below stay zero, as the reference's zeroed work arrays read there
    HIPCHK(hipMemset(s.adfh, 0, ELn * kb * cb));
    HIPCHK(hipMemset(s.LO, 0, NLn * kb * cb));
    HIPCHK(hipMemset(s.adfv, 0, NIn * kb * cb));
    // tracer blocks (unused lanes of the last block stay zero and are never advanced)
    const int nb = n_blocks(s);
    s.dT.assign(nb, nullptr);
    s.tabbuf.assign(nb, nullptr);
    s.tab_on.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        if (dalloc_bytes(s, &s.dT[b], NLn * kb * sb)) return -1;
        HIPCHK(hipMemset(s.dT[b], 0, NLn * kb * sb));
    }
    s.dR.assign(nb, nullptr);
    s.LOe = nullptr;
    if (s.compensated && s.prec != PREC_F64) {
        for (int b = 0; b < nb; ++b) {  // FP32 remainders whatever the storage type
            if (dalloc_bytes(s, &s.dR[b], NLn * kb * sizeof(float))) return -1;
            HIPCHK(hipMemset(s.dR[b], 0, NLn * kb * sizeof(float)));
        }
        if (dalloc_bytes(s, &s.LOe, NLn * kb * cb)) return -1;
        HIPCHK(hipMemset(s.LOe, 0, NLn * kb * cb));
    }
    if (build_halo_tiles(s) || build_edge_tiles(s)) return -1;
    HIPCHK(hipDeviceSynchronize());
    s.on_device = true;
    s.gen++;
    return 0;
}

// The remainder buffer of a tracer block's storage (compensated storage), else null.
void *resid_of(const State &s, const void *buf) {
    for (size_t b = 0; b < s.dT.size(); ++b)
        if (s.dT[b] == buf) return b < s.dR.size() ? s.dR[b] : nullptr;
    return nullptr;
}

// reference-layout host values -> lane of a device tracer block (through the staging buffer); a
// block with compensated storage also gets the rounding remainders
int to_device(State &s, const double *host_or_dev, bool is_device, void *dst, int lane) {
    void *res = resid_of(s, dst);
    const int L = s.nl - 1;
    const double *src = host_or_dev;
    if (!is_device) {
        HIPCHK(hipMemcpyAsync(s.stage, host_or_dev, sizeof(double) * (size_t)L * s.N, hipMemcpyHostToDevice, s.stream));
        src = s.stage;
    }
    hipError_t e;
    switch (s.prec) {
    case PREC_F64: e = launch_relayout<double>(src, (double *)dst, nullptr, nullptr, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 0, s.stream); break;
    case PREC_F32: e = launch_relayout<float>(src, (float *)dst, nullptr, nullptr, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 0, s.stream, (float *)res); break;
    case PREC_F32_F16: e = launch_relayout<f16_t>(src, (f16_t *)dst, nullptr, nullptr, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 0, s.stream, (float *)res); break;
    default: e = launch_relayout<bf16_t>(src, (bf16_t *)dst, nullptr, nullptr, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 0, s.stream, (float *)res); break;
    }
    if (e != hipSuccess) return fail("relayout launch failed: %s", hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(s.stream));
    return 0;
}

// lane of a device tracer block -> reference layout (host or device destination), stored value
// plus remainder with compensated storage
int from_device(State &s, const void *src, int lane, double *out, bool out_is_device) {
    void *res = resid_of(s, src);
    const int L = s.nl - 1;
    double *dst = out_is_device ? out : s.stage;
    hipError_t e;
    switch (s.prec) {
    case PREC_F64: e = launch_relayout<double>(nullptr, nullptr, (const double *)src, dst, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 1, s.stream); break;
    case PREC_F32: e = launch_relayout<float>(nullptr, nullptr, (const float *)src, dst, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 1, s.stream, (float *)res); break;
    case PREC_F32_F16: e = launch_relayout<f16_t>(nullptr, nullptr, (const f16_t *)src, dst, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 1, s.stream, (float *)res); break;
    default: e = launch_relayout<bf16_t>(nullptr, nullptr, (const bf16_t *)src, dst, s.d_perm_n, s.N, L, s.Np, s.kb, lane, 1, s.stream, (float *)res); break;
    }
    if (e != hipSuccess) return fail("relayout launch failed: %s", hipGetErrorString(e));
    if (!out_is_device)
        HIPCHK(hipMemcpyAsync(out, s.stage, sizeof(double) * (size_t)L * s.N, hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    return 0;
}

// valuesAB (and valuesold) storage of every block for ab_mode 1 / 2, set to the values, or to
// the host copies kept across a rebuild where there are any.
int ab_prepare(State &s) {
    if (s.ab_mode == 0) return 0;
    const size_t bytes = (size_t)(s.nl - 1) * s.Np * s.kb * storage_bytes(s.prec);
    if (s.toldbuf.size() != s.dT.size()) s.toldbuf.assign(s.dT.size(), nullptr);
    for (size_t b = 0; b < s.dT.size(); ++b) {
        if (!s.tabbuf[b] && dalloc_bytes(s, &s.tabbuf[b], bytes)) return -1;
        HIPCHK(hipMemcpyAsync(s.tabbuf[b], s.dT[b], bytes, hipMemcpyDeviceToDevice, s.stream));
        if (s.ab_mode == 2) {
            if (s.tabbuf2.size() != s.dT.size()) s.tabbuf2.assign(s.dT.size(), nullptr);
            if (!s.toldbuf[b] && dalloc_bytes(s, &s.toldbuf[b], bytes)) return -1;
            if (!s.tabbuf2[b] && dalloc_bytes(s, &s.tabbuf2[b], bytes)) return -1;
            HIPCHK(hipMemcpyAsync(s.toldbuf[b], s.dT[b], bytes, hipMemcpyDeviceToDevice, s.stream));
        }
        s.tab_on[b] = 1;
    }
    for (int i = 0; i < s.ntr; ++i) {
        const int b = i / s.kb, lane = i % s.kb;
        if (i < (int)s.ab_host.size() && !s.ab_host[i].empty() &&
            to_device(s, s.ab_host[i].data(), false, s.tabbuf[b], lane))
            return -1;
        if (s.ab_mode == 2 && i < (int)s.old_host.size() && !s.old_host[i].empty() &&
            to_device(s, s.old_host[i].data(), false, s.toldbuf[b], lane))
            return -1;
    }
    s.ab_host.clear();
    s.old_host.clear();
    s.ab_fresh = true;
    HIPCHK(hipStreamSynchronize(s.stream));
    return 0;
}

int ensure_device(State &s) {
    if (!s.initialized) return fail("tracer not initialized. Call tracer_init_complete first.");
    if (s.on_device) return 0;
    if (build_device(s)) return -1;
    for (int i = 0; i < s.ntr; ++i)
        if (to_device(s, s.values[i].data(), false, s.dT[i / s.kb], i % s.kb)) return -1;
    return ab_prepare(s);
}

// Values of every tracer back to the host copies (before the device storage is rebuilt), with
// valuesAB / valuesold when the time-level mode keeps them.
int pull_values(State &s) {
    if (!s.on_device) return 0;
    for (int i = 0; i < s.ntr; ++i)
        if (from_device(s, s.dT[i / s.kb], i % s.kb, s.values[i].data(), false)) return -1;
    s.ab_host.clear();
    s.old_host.clear();
    if (s.ab_mode != 0) {
        const size_t n = (size_t)(s.nl - 1) * s.N;
        s.ab_host.assign(s.ntr, std::vector<double>(n));
        if (s.ab_mode == 2) s.old_host.assign(s.ntr, std::vector<double>(n));
        for (int i = 0; i < s.ntr; ++i) {
            const int b = i / s.kb, lane = i % s.kb;
            if (s.tab_on[b] && from_device(s, s.tabbuf[b], lane, s.ab_host[i].data(), false)) return -1;
            if (!s.tab_on[b]) s.ab_host[i] = s.values[i];
            if (s.ab_mode == 2 && s.toldbuf[b] && from_device(s, s.toldbuf[b], lane, s.old_host[i].data(), false))
                return -1;
        }
    }
    return 0;
}

int ensure_dropin_buffers(State &s) {
    const size_t NLn = (size_t)(s.nl - 1) * s.Np, ELn = (size_t)(s.nl - 1) * s.Ep, NIn = (size_t)s.nl * s.Np;
    const size_t cb = compute_bytes(s.prec);
    if (!s.acc_h) {
        if (dalloc(s, &s.acc_h, NLn) || dalloc(s, &s.acc_v, NLn) || dalloc_bytes(s, &s.dbg_h, ELn * cb) ||
            dalloc_bytes(s, &s.dbg_v, NIn * cb))
            return -1;
        HIPCHK(hipMemsetAsync(s.acc_h, 0, NLn * sizeof(double), s.stream));
        HIPCHK(hipMemsetAsync(s.acc_v, 0, NLn * sizeof(double), s.stream));
        // the limited fluxes are written at the levels an edge / node has; the rest stays zero
        HIPCHK(hipMemsetAsync(s.dbg_h, 0, ELn * cb, s.stream));
        HIPCHK(hipMemsetAsync(s.dbg_v, 0, NIn * cb, s.stream));
    }
    return 0;
}

// ab_parity (AB2 steps): 0 reads valuesAB from the primary buffer and writes the next step's to the
// second one, 1 the other way round.
template <class S_, class C_>
StepPtrs<S_, C_> make_ptrs(State &s, int blk, unsigned long long mask, bool dropin, uint32_t seed, int ab_parity) {
    StepPtrs<S_, C_> p{};
    p.tid0 = blk * s.kb;
    p.seed = (s.prec != PREC_F64 && s.stochastic) ? seed : 0u;
    p.seed_dev = p.seed ? s.d_seed : nullptr;
    p.T = static_cast<S_ *>(s.dT[blk]);
    p.TAB = static_cast<const S_ *>(s.tab_on[blk] ? s.tabbuf[blk] : s.dT[blk]);
    if (s.ab_mode == 2 && !dropin && blk < (int)s.tabbuf2.size() && s.tabbuf2[blk]) {
        p.TAB = static_cast<const S_ *>(ab_parity ? s.tabbuf2[blk] : s.tabbuf[blk]);
        p.ab_next = static_cast<S_ *>(ab_parity ? s.tabbuf[blk] : s.tabbuf2[blk]);
        p.ab_old = static_cast<S_ *>(s.toldbuf[blk]);
        p.ab_c0 = 0.5 + s.ab_eps;
        p.ab_c1 = 1.5 + s.ab_eps;
    }
    p.adfh = static_cast<C_ *>(s.adfh);
    p.LO = static_cast<C_ *>(s.LO);
    p.LOe = static_cast<C_ *>(s.LOe);
    p.Rp = static_cast<C_ *>(s.Rp);
    p.Rm = static_cast<C_ *>(s.Rm);
    p.adfv = static_cast<C_ *>(s.adfv);
    p.mask = mask;
    p.tri = static_cast<double2 *>(s.tri);
    p.cqv = static_cast<C_ *>(s.cqv);
    p.resid = dropin ? nullptr : static_cast<float *>(resid_of(s, s.dT[blk]));
    p.lane_mem = s.nmem > 1 ? s.d_lane_mem + (size_t)blk * s.kb : nullptr;
    p.mg = static_cast<C_ *>(s.mg);
    if (dropin) {
        p.acc_h = s.acc_h;
        p.acc_v = s.acc_v;
        p.acc_init = s.acc_zero ? 1 : 0;
        if (s.keep_work) {
            p.dbg_h = static_cast<C_ *>(s.dbg_h);
            p.dbg_v = static_cast<C_ *>(s.dbg_v);
        }
    }
    return p;
}

// The launch configuration of a pass: the aux stream only while concurrency is on.
LaunchCfg launch_cfg(const State &s) {
    LaunchCfg c = s.cfg;
    if (!s.concurrent) c.aux = nullptr;
    return c;
}

// tin / tout (fused limiter steps): the block's values are read from tin and the update written to
// tout (the two buffers of the block, alternating by step); null: the primary buffer in place.
hipError_t launch_block(State &s, int blk, unsigned long long mask, const SchemeSel &sc, double dt, bool dropin,
                        hipEvent_t *ev, uint32_t seed = 0, int ab_parity = 0, void *tin = nullptr,
                        void *tout = nullptr) {
    // cavity meshes: the limiter reads edge gradients filled from this tracer's valuesAB where the
    // scheme fills them (MUSCL / MFCT) and always for the drop-in, whose caller fills them
    SchemeArgs a{sc.hor, sc.ver, sc.fct, sc.ph, sc.pv, dt, s.wsplit && sc.fct, sc.hor != HOR_UPW1 || dropin};
    if (tin) {
        if (s.prec != PREC_F64) return hipErrorInvalidValue;
        StepPtrs<double, double> p = make_ptrs<double, double>(s, blk, mask, dropin, seed, ab_parity);
        if (!s.tab_on[blk]) p.TAB = static_cast<const double *>(tin);
        p.T = static_cast<double *>(tin);
        p.Tout = static_cast<double *>(tout);
        return launch_step<double, double, double>(s.m64, p, s.kb, a, launch_cfg(s), s.stream, ev);
    }
    switch (s.prec) {
    case PREC_F64: return launch_step<double, double, double>(s.m64, make_ptrs<double, double>(s, blk, mask, dropin, seed, ab_parity), s.kb, a, launch_cfg(s), s.stream, ev);
    case PREC_F32: return launch_step<float, double, float>(s.m64, make_ptrs<float, float>(s, blk, mask, dropin, seed, ab_parity), s.kb, a, launch_cfg(s), s.stream, ev);
    case PREC_F32_F16: return launch_step<f16_t, double, float>(s.m64, make_ptrs<f16_t, float>(s, blk, mask, dropin, seed, ab_parity), s.kb, a, launch_cfg(s), s.stream, ev);
    default: return launch_step<bf16_t, double, float>(s.m64, make_ptrs<bf16_t, float>(s, blk, mask, dropin, seed, ab_parity), s.kb, a, launch_cfg(s), s.stream, ev);
    }
}

// Tiles of the fused path: consecutive device nodes (a stretch of the Hilbert curve) as owned
// nodes, plus three rings of neighbours; shrunk until the region, its edge entries and its
// elements fit the kernel's thread and LDS budget.  Tables are packed as the kernel reads them
// (fesomtra_tile.hpp).  Returns -1 (with a message) if the mesh exceeds the packing limits.
int build_tiles(State &s) {
    const int N = s.N, A = s.max_adj;
    if (s.nl > 128) return fail("fused step: nl=%d above 128", s.nl);
    if ((int64_t)s.E * 2 >= (1 << 24)) return fail("fused step: %d edges above the packing limit", s.E);
    std::vector<int> stamp(N, -1), loc(N, 0), estamp(s.T, -1), eloc(s.T, 0);
    std::vector<int> t_slot0, t_nslot, t_sp0, t_ent0, t_nent, t_sel0, t_nsel, t_el0, t_nel, t_lmax;
    std::vector<int> s_node, s_ring, s_eptr, s_elptr, sel, el_id, el_loc;
    std::vector<int2> ent;
    int start = 0, stamp_id = 0;
    std::vector<int> region, ring_of, tel;
    while (start < N) {
        int P = std::min(kTileOwned, N - start);
        for (;;) {
            ++stamp_id;
            region.clear();
            ring_of.clear();
            for (int n = start; n < start + P; ++n) {
                stamp[n] = stamp_id;
                loc[n] = (int)region.size();
                region.push_back(n);
                ring_of.push_back(0);
            }
            size_t f0 = 0;
            for (int r = 1; r <= 3; ++r) {
                const size_t f1 = region.size();
                for (size_t i = f0; i < f1; ++i) {
                    const int u = region[i];
                    for (int k = s.h_ptr[u]; k < s.h_ptr[u + 1]; ++k) {
                        const int v = s.h_oth[k];
                        if (stamp[v] != stamp_id) {
                            stamp[v] = stamp_id;
                            loc[v] = (int)region.size();
                            region.push_back(v);
                            ring_of.push_back(r);
                        }
                    }
                }
                f0 = f1;
            }
            bool ok = (int)region.size() <= kTileThreads;
            tel.clear();
            int nent = 0, nsel = 0;
            if (ok) {
                auto add_el = [&](int el) {
                    if (estamp[el] != stamp_id) {
                        estamp[el] = stamp_id;
                        eloc[el] = (int)tel.size();
                        tel.push_back(el);
                    }
                };
                for (size_t i = 0; i < region.size(); ++i) {
                    if (ring_of[i] > 2) continue;
                    const int u = region[i];
                    nent += s.h_ptr[u + 1] - s.h_ptr[u];
                    nsel += s.h_nie_cnt[u];
                    for (int k = 0; k < s.h_nie_cnt[u]; ++k) add_el(s.h_nie[(size_t)u * A + k]);
                    if (ring_of[i] <= 1) {
                        if (s.h_ptr[u + 1] - s.h_ptr[u] > kTileDeg)
                            return fail("fused step: node with %d edges (more than %d)", s.h_ptr[u + 1] - s.h_ptr[u], kTileDeg);
                        for (int k = s.h_ptr[u]; k < s.h_ptr[u + 1]; ++k) {
                            const int2 ud = s.h_ud[s.h_e2[k] >> 1];
                            if (ud.x >= 0 && ud.y >= 0) {
                                add_el(ud.x);
                                add_el(ud.y);
                            }
                        }
                    }
                }
                ok = (int)tel.size() <= kTileElems && nent <= kTileEntries && nsel <= kTileSel;
                for (size_t i = 0; i < tel.size() && ok; ++i)
                    for (int c = 0; c < 3; ++c)
                        if (stamp[s.h_en[3 * (size_t)tel[i] + c]] != stamp_id) ok = false;  // node outside the region
            }
            if (ok) break;
            if (P == 1) return fail("fused step: a one-node tile does not fit the region limits");
            P = std::max(1, P * 3 / 4);
        }
        t_slot0.push_back((int)s_node.size());
        t_nslot.push_back((int)region.size());
        t_sp0.push_back((int)s_eptr.size());
        t_ent0.push_back((int)ent.size());
        t_sel0.push_back((int)sel.size());
        t_el0.push_back((int)el_id.size());
        t_nel.push_back((int)tel.size());
        const int ent_base = (int)ent.size(), sel_base = (int)sel.size();
        int lmax = 0;
        for (size_t i = 0; i < region.size(); ++i) {
            const int u = region[i];
            s_node.push_back(u);
            s_ring.push_back(ring_of[i]);
            s_eptr.push_back((int)ent.size() - ent_base);
            s_elptr.push_back((int)sel.size() - sel_base);
            lmax = std::max(lmax, s.h_nlev[u] - 1);
            if (ring_of[i] > 2) continue;
            for (int k = s.h_ptr[u]; k < s.h_ptr[u + 1]; ++k) {
                const int2 ud = s.h_ud[s.h_e2[k] >> 1];
                const bool has_ud = ring_of[i] <= 1 && ud.x >= 0 && ud.y >= 0;
                const int up = has_ud ? eloc[ud.x] : 2047, dn = has_ud ? eloc[ud.y] : 2047;
                ent.push_back(make_int2((s.h_e2[k] << 7) | s.h_top[k], loc[s.h_oth[k]] | (up << 10) | (dn << 21)));
            }
            for (int k = 0; k < s.h_nie_cnt[u]; ++k) sel.push_back(eloc[s.h_nie[(size_t)u * A + k]]);
        }
        s_eptr.push_back((int)ent.size() - ent_base);
        s_elptr.push_back((int)sel.size() - sel_base);
        t_nent.push_back((int)ent.size() - ent_base);
        t_nsel.push_back((int)sel.size() - sel_base);
        for (const int el : tel) {
            el_id.push_back(el);
            el_loc.push_back(loc[s.h_en[3 * (size_t)el]] | (loc[s.h_en[3 * (size_t)el + 1]] << 10) |
                             (loc[s.h_en[3 * (size_t)el + 2]] << 20));
        }
        t_lmax.push_back(lmax);
        start += P;
    }
    while (t_slot0.size() % kXcds) {  // empty tiles pad the grid to the XCD count
        t_slot0.push_back((int)s_node.size());
        t_nslot.push_back(0);
        t_sp0.push_back((int)s_eptr.size());
        s_eptr.push_back(0);
        s_elptr.push_back(0);
        t_ent0.push_back((int)ent.size());
        t_nent.push_back(0);
        t_sel0.push_back((int)sel.size());
        t_nsel.push_back(0);
        t_el0.push_back((int)el_id.size());
        t_nel.push_back(0);
        t_lmax.push_back(0);
    }
    TileMesh &t = s.tm;
    t.ntiles = (int)t_slot0.size();
    if (upload(s, &t.t_slot0, t_slot0) || upload(s, &t.t_nslot, t_nslot) || upload(s, &t.t_sp0, t_sp0) ||
        upload(s, &t.t_ent0, t_ent0) || upload(s, &t.t_nent, t_nent) || upload(s, &t.t_sel0, t_sel0) ||
        upload(s, &t.t_nsel, t_nsel) || upload(s, &t.t_el0, t_el0) || upload(s, &t.t_nel, t_nel) ||
        upload(s, &t.t_lmax, t_lmax) || upload(s, &t.s_node, s_node) || upload(s, &t.s_ring, s_ring) ||
        upload(s, &t.s_eptr, s_eptr) || upload(s, &t.s_elptr, s_elptr) || upload(s, &t.ent, ent) ||
        upload(s, &t.sel, sel) || upload(s, &t.el_id, el_id) || upload(s, &t.el_loc, el_loc))
        return -1;
    s.tiles_built = true;
    return 0;
}

// One launch sequence per (block, scheme): the lanes of a block that share a scheme.
struct Launch {
    int blk;
    unsigned long long mask;
    SchemeSel sc;
};

// The fused limiter (k_limit) takes the bounds, ratios and update of every pass of a step: FP64,
// node tiles with two-ring halos built for the block width, the FCT limiter on every pass, no
// cavities, per-node time steps, ensembles or AB2 time levels.  The update then goes to the
// block's second buffer (steps alternate between the two).
bool limiter_applies(const State &s, const std::vector<Launch> &plan) {
    if (!s.fused_limiter || s.m64.rt_n < 1 || s.m64.ht_kb != s.kb || s.kb > 4 || s.prec != PREC_F64) return false;
    if (!s.node_dt.empty() || s.nmem > 1 || s.m64.ulev_n || s.ab_mode == 2 || s.fused) return false;
    for (const Launch &l : plan)
        if (!l.sc.fct) return false;
    return !plan.empty();
}

// The fused path applies to FP64 blocks of one tracer with the FCT limiter (without the
// implicit vertical split, with valuesAB := values).
bool fused_applies(const State &s, const SchemeSel &sc) {
    if (!s.node_dt.empty()) return false;  // the fused kernel takes one time step for all nodes
    if (s.nmem > 1) return false;          // and one velocity
    if (s.m64.ulev_n) return false;        // and has no cavity levels
    return s.fused && !s.wsplit && s.ab_mode == 0 && s.prec == PREC_F64 && s.kb == 1 && sc.fct;
}

// Scratch of the implicit vertical solve: one (c', t') pair per layer and lane.
int ensure_wsplit(State &s) {
    if (!s.wsplit || s.tri) return 0;
    return dalloc_bytes(s, &s.tri, (size_t)(s.nl - 1) * s.Np * s.kb * sizeof(double2));
}

// Second value buffers of the blocks (output of fused steps), with the values below the bottom
// copied once from the primary ones.
int ensure_t2(State &s) {
    const size_t bytes = (size_t)(s.nl - 1) * s.Np * s.kb * storage_bytes(s.prec);
    if (s.dT2.size() != s.dT.size()) {
        s.dT2.assign(s.dT.size(), nullptr);
        s.t2_sync.assign(s.dT.size(), 0);
    }
    for (size_t b = 0; b < s.dT.size(); ++b) {
        if (!s.dT2[b] && dalloc_bytes(s, &s.dT2[b], bytes)) return -1;
        if (!s.t2_sync[b]) {
            HIPCHK(hipMemcpyAsync(s.dT2[b], s.dT[b], bytes, hipMemcpyDeviceToDevice, s.stream));
            s.t2_sync[b] = 1;
        }
    }
    return 0;
}

int ensure_fused(State &s) {
    if (!s.tiles_built && build_tiles(s)) return -1;
    return ensure_t2(s);
}

// Marks the second buffers stale after the primary storage was written from outside a step.
void touch_values(State &s) { std::fill(s.t2_sync.begin(), s.t2_sync.end(), 0); }

uint32_t next_seed(State &s) {
    if (++s.step_seed == 0u) s.step_seed = 1u;
    return s.step_seed;
}
// The stochastic rounding's step counter is used (single precision with stochastic rounding).
bool seeded(const State &s) { return s.prec != PREC_F64 && s.stochastic; }

// valuesAB / valuesold of a block from its values (FESOM2's AB2 interpolation, before a step).
hipError_t launch_ab2_block(State &s, int b) {
    const size_t n = (size_t)(s.nl - 1) * s.Np * s.kb;
    switch (s.prec) {
    case PREC_F64: return launch_ab2<double>((const double *)s.dT[b], (double *)s.tabbuf[b], (double *)s.toldbuf[b], n, s.ab_eps, s.stream);
    case PREC_F32: return launch_ab2<float>((const float *)s.dT[b], (float *)s.tabbuf[b], (float *)s.toldbuf[b], n, s.ab_eps, s.stream);
    case PREC_F32_F16: return launch_ab2<f16_t>((const f16_t *)s.dT[b], (f16_t *)s.tabbuf[b], (f16_t *)s.toldbuf[b], n, s.ab_eps, s.stream);
    default: return launch_ab2<bf16_t>((const bf16_t *)s.dT[b], (bf16_t *)s.tabbuf[b], (bf16_t *)s.toldbuf[b], n, s.ab_eps, s.stream);
    }
}

std::vector<Launch> launch_plan(const State &s) {
    std::vector<Launch> plan;
    for (int b = 0; b < n_blocks(s); ++b) {
        const size_t first = plan.size();
        for (int l = 0; l < s.kb && b * s.kb + l < s.ntr; ++l) {
            const SchemeSel &sc = s.schemes[b * s.kb + l];
            size_t i = first;
            while (i < plan.size() && !(plan[i].sc == sc)) ++i;
            if (i == plan.size()) plan.push_back(Launch{b, 0ull, sc});
            plan[i].mask |= 1ull << l;
        }
    }
    return plan;
}

// Events of the kernel timing per step: k_qflux start and end, then 5 per pass (around its 4
// stages); the last pass's end closes the step.
size_t step_events(size_t passes) { return 2 + 5 * passes; }

// nsteps of the launch plan on the library's stream.  Fused blocks alternate between their two
// buffers and end in the primary one (one device copy after an odd number of steps), so the
// stored values are always in dT whatever the path.  evs (optional) receives step_events(passes)
// events per step.  With stochastic rounding every step first advances the device step counter
// (d_seed), which the caller set to step_seed before the sequence (also before a graph replay).
// AB2 (ab_mode 2): the update of every step also writes the next step's valuesAB and valuesold
// (k_final); valuesAB alternates between two buffers and ends in the primary one.  The
// interpolation kernel runs before the first step when valuesAB is stale (values or the mode
// were just set) and always at the start of a captured sequence (it is idempotent: valuesAB
// then already equals the interpolation of values and valuesold).
hipError_t run_steps(State &s, const std::vector<Launch> &plan, int nsteps, double dt, hipEvent_t *evs,
                     bool capture = false) {
    std::vector<char> where(s.dT.size(), 0);
    const bool ab2 = s.ab_mode == 2;
    const bool lim = limiter_applies(s, plan);  // every pass through k_limit, values alternating
    const size_t nev = step_events(plan.size());
    for (int st = 0; st < nsteps; ++st) {
        hipEvent_t *sev = evs ? &evs[(size_t)st * nev] : nullptr;
        hipError_t e = sev ? hipEventRecord(sev[0], s.stream) : hipSuccess;
        // edge transport of the velocity, shared by every tracer pass of the step
        if (e == hipSuccess) e = launch_qflux(s.m64, s.stream);
        if (e == hipSuccess && sev) e = hipEventRecord(sev[1], s.stream);
        if (e != hipSuccess) return e;
        s.q_vel = s.vel_gen;
        if (ab2 && st == 0 && (s.ab_fresh || capture))
            for (size_t b = 0; b < s.dT.size() && e == hipSuccess; ++b) e = launch_ab2_block(s, (int)b);
        if (e == hipSuccess && seeded(s)) e = launch_seed(s.d_seed, 0u, 1, s.stream);
        if (e != hipSuccess) return e;
        const uint32_t seed = 1u;  // stochastic rounding on; the step's value is read from d_seed
        for (size_t gi = 0; gi < plan.size(); ++gi) {
            const Launch &l = plan[gi];
            hipEvent_t *ev = evs ? &sev[2 + 5 * gi] : nullptr;
            if (fused_applies(s, l.sc)) {
                double *in = static_cast<double *>(where[l.blk] ? s.dT2[l.blk] : s.dT[l.blk]);
                double *out = static_cast<double *>(where[l.blk] ? s.dT[l.blk] : s.dT2[l.blk]);
                for (int k = 0; k < 4 && ev; ++k)
                    if ((e = hipEventRecord(ev[k], s.stream)) != hipSuccess) return e;
                e = launch_tile_step(s.m64, s.tm, in, out, l.sc.hor, l.sc.ver, l.sc.ph, l.sc.pv, dt, s.stream);
                if (e == hipSuccess && ev) e = hipEventRecord(ev[4], s.stream);
                where[l.blk] ^= 1;
            } else if (lim) {
                // every pass of the block in this step reads one buffer and writes the other
                void *in = where[l.blk] ? s.dT2[l.blk] : s.dT[l.blk];
                void *out = where[l.blk] ? s.dT[l.blk] : s.dT2[l.blk];
                e = launch_block(s, l.blk, l.mask, l.sc, dt, false, ev, seed, 0, in, out);
            } else {
                e = launch_block(s, l.blk, l.mask, l.sc, dt, false, ev, seed, ab2 ? (st & 1) : 0);
            }
            if (e != hipSuccess) return e;
        }
        if (lim)
            for (size_t b = 0; b < where.size(); ++b) where[b] ^= 1;
    }
    const size_t bytes = (size_t)(s.nl - 1) * s.Np * s.kb * storage_bytes(s.prec);
    for (size_t b = 0; b < where.size(); ++b)
        if (where[b]) {
            hipError_t e = hipMemcpyAsync(s.dT[b], s.dT2[b], bytes, hipMemcpyDeviceToDevice, s.stream);
            if (e != hipSuccess) return e;
        }
    if (ab2 && (nsteps & 1))  // the next step's valuesAB was written to the second buffer
        for (size_t b = 0; b < s.dT.size(); ++b) {
            hipError_t e = hipMemcpyAsync(s.tabbuf[b], s.tabbuf2[b], bytes, hipMemcpyDeviceToDevice, s.stream);
            if (e != hipSuccess) return e;
        }
    if (!capture && nsteps > 0) s.ab_fresh = false;
    return hipSuccess;
}

// Captures nsteps of the launch plan into one executable graph (kept until the state changes).
int capture_graph(State &s, const std::vector<Launch> &plan, int nsteps, double dt) {
    drop_graph(s);
    HIPCHK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
    const hipError_t err = run_steps(s, plan, nsteps, dt, nullptr, true);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s.stream, &graph);
    if (err != hipSuccess || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail("graph capture failed: %s", hipGetErrorString(err != hipSuccess ? err : ec));
    }
    const hipError_t ei = hipGraphInstantiate(&s.gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        s.gexec = nullptr;
        return fail("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    }
    s.g_nsteps = nsteps;
    s.g_dt = dt;
    s.g_gen = s.gen;
    return 0;
}

int advect(State &s, int nsteps, double dt) {
    if (ensure_device(s)) return -1;
    if (s.ntr == 0) return fail("no tracers allocated (tracer_allocate_tracers)");
    if (ensure_wsplit(s)) return -1;
    if (s.ab_mode == 0) std::fill(s.tab_on.begin(), s.tab_on.end(), 0);  // valuesAB := values
    const std::vector<Launch> plan = launch_plan(s);
    if (s.m64.cq_n > 0) {
        // the limiter of a UPW1 tracer reads the gradients another tracer's step left in the
        // reference's shared work array: defined here only when no tracer fills them
        bool fills = false, reads = false;
        for (const SchemeSel &sc : s.schemes) {
            fills = fills || sc.hor != HOR_UPW1;
            reads = reads || (sc.fct && sc.hor == HOR_UPW1);
        }
        if (fills && reads)
            return fail("cavity mesh: a UPW1 tracer with the FCT limiter next to MUSCL / MFCT tracers reads edge "
                        "gradients of another tracer in the reference (not reproduced); advance them separately");
    }
    if (s.fused && s.prec == PREC_F64 && s.kb == 1 && ensure_fused(s)) return -1;
    if (limiter_applies(s, plan) && ensure_t2(s)) return -1;
    if (!plan.empty()) s.last_fct = plan.back().sc.fct;
    if (nsteps > 0) s.dropin_valid = false;  // the work arrays now hold this call's intermediates
    for (const Launch &l : plan)
        if (nsteps > 0 && !l.sc.fct && s.prec != PREC_F64) s.lo_dirty = true;
    // the device step counter starts from the host's; the host mirrors the steps' advances
    if (seeded(s) && nsteps > 0) {
        const hipError_t e = launch_seed(s.d_seed, s.step_seed, 0, s.stream);
        if (e != hipSuccess) return fail("kernel launch failed: %s", hipGetErrorString(e));
        for (int i = 0; i < nsteps; ++i) next_seed(s);
    }
    if (s.use_graph && !s.kernel_timing && nsteps > 0) {
        if (!s.gexec || s.g_nsteps != nsteps || s.g_dt != dt || s.g_gen != s.gen)
            if (capture_graph(s, plan, nsteps, dt)) return -1;
        HIPCHK(hipEventRecord(s.ev0, s.stream));
        HIPCHK(hipGraphLaunch(s.gexec, s.stream));
        if (nsteps > 0) s.ab_fresh = false;
        HIPCHK(hipEventRecord(s.ev1, s.stream));
        HIPCHK(hipStreamSynchronize(s.stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
        s.last_ms = ms;
        return 0;
    }
    std::vector<hipEvent_t> evs;
    const size_t nev = step_events(plan.size());
    if (s.kernel_timing) {
        evs.resize((size_t)nsteps * nev);
        for (auto &e : evs) HIPCHK(hipEventCreate(&e));
    }
    HIPCHK(hipEventRecord(s.ev0, s.stream));
    {
        const hipError_t e = run_steps(s, plan, nsteps, dt, s.kernel_timing ? evs.data() : nullptr);
        if (e != hipSuccess) return fail("kernel launch failed: %s", hipGetErrorString(e));
    }
    HIPCHK(hipEventRecord(s.ev1, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    s.last_ms = ms;
    if (s.kernel_timing && !plan.empty()) {
        for (int st = 0; st < nsteps; ++st) {
            hipEvent_t *sev = &evs[(size_t)st * nev];
            float d = 0.f;
            HIPCHK(hipEventElapsedTime(&d, sev[0], sev[1]));
            s.qflux_ms += d;
            HIPCHK(hipEventElapsedTime(&d, sev[0], sev[nev - 1]));
            s.step_ms += d;
            for (size_t gi = 0; gi < plan.size(); ++gi)
                for (int k = 0; k < 4; ++k) {
                    HIPCHK(hipEventElapsedTime(&d, sev[2 + 5 * gi + k], sev[2 + 5 * gi + k + 1]));
                    s.stage_ms[k] += d;
                }
        }
        s.stage_count += (long)nsteps * (long)plan.size();
        s.step_count += nsteps;
    }
    for (auto &e : evs) (void)hipEventDestroy(e);
    return 0;
}

// Lane `lane` of a device work array [ld][pad][kb] (device order, element size cb) -> reference
// layout host.
int download_work(State &s, const void *d, int ld, size_t pad, const std::vector<int> &perm, double *out,
                  size_t cb, int kb, int lane) {
    std::vector<char> h((size_t)ld * pad * cb);
    HIPCHK(hipStreamSynchronize(s.stream));
    HIPCHK(hipMemcpy2D(h.data(), cb, static_cast<const char *>(d) + (size_t)lane * cb, (size_t)kb * cb, cb,
                       (size_t)ld * pad, hipMemcpyDeviceToHost));
    for (size_t r = 0; r < perm.size(); ++r)
        for (int k = 0; k < ld; ++k) {
            const size_t i = (size_t)k * pad + r;
            out[(size_t)perm[r] * ld + k] = cb == 8 ? ((const double *)h.data())[i] : (double)((const float *)h.data())[i];
        }
    return 0;
}

int stats(State &s, int id, int all_levels, double *mn, double *mx, double *sm, double *vs) {
    if (ensure_device(s)) return -1;
    if (id < 1 || id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", id, s.ntr);
    const void *T = s.dT[(id - 1) / s.kb];
    const void *R = resid_of(s, T);
    const int kb = s.kb, lane = (id - 1) % s.kb;
    hipError_t e;
    switch (s.prec) {
    case PREC_F64: e = launch_stats<double, double>(s.m64, (const double *)T, kb, lane, all_levels, s.part, kStatsBlocks, s.stream); break;
    case PREC_F32: e = launch_stats<float, double>(s.m64, (const float *)T, kb, lane, all_levels, s.part, kStatsBlocks, s.stream, (const float *)R); break;
    case PREC_F32_F16: e = launch_stats<f16_t, double>(s.m64, (const f16_t *)T, kb, lane, all_levels, s.part, kStatsBlocks, s.stream, (const float *)R); break;
    default: e = launch_stats<bf16_t, double>(s.m64, (const bf16_t *)T, kb, lane, all_levels, s.part, kStatsBlocks, s.stream, (const float *)R); break;
    }
    if (e != hipSuccess) return fail("stats launch failed: %s", hipGetErrorString(e));
    std::vector<double4> h(kStatsBlocks);
    HIPCHK(hipStreamSynchronize(s.stream));
    HIPCHK(hipMemcpy(h.data(), s.part, sizeof(double4) * kStatsBlocks, hipMemcpyDeviceToHost));
    double a = INFINITY, b = -INFINITY, c = 0.0, d = 0.0;
    for (const auto &p : h) {
        a = std::min(a, p.x);
        b = std::max(b, p.y);
        c += p.z;
        d += p.w;
    }
    if (mn) *mn = a;
    if (mx) *mx = b;
    if (sm) *sm = c;
    if (vs) *vs = d;
    return 0;
}


// Library state -> restart items (lib/io_restart_derivedtype.F90).  A state read from restart
// files keeps every item of those files; the tracers, schemes, MUSCL support arrays and
// velocities are written from the current state, the other items of a state built from arrays
// are what the mesh setters hold (connectivity and levels 1-based as in the reference).
void restart_of_state(const State &s, RestartFiles &o) {
    if (s.from_restart) o = s.rs;
    const int N = s.N, T = s.T, E = s.E, nl = s.nl, L = nl - 1;
    auto plus1 = [](const std::vector<int> &v) {
        std::vector<int> r(v.size());
        for (size_t i = 0; i < v.size(); ++i) r[i] = v[i] + 1;  // -1 (none) becomes 0
        return r;
    };
    if (!s.from_restart) {
        o.set_scalar_int("mesh.nod2D", N);
        o.set_scalar_int("mesh.edge2D", E);
        o.set_scalar_int("mesh.edge2D_in", E);
        o.set_scalar_int("mesh.elem2D", T);
        o.set_scalar_int("mesh.nl", nl);
        o.set_ints("mesh.elem2D_nodes", plus1(s.enodes), {3, T});
        o.set_ints("mesh.edges", plus1(s.edges), {2, E});
        o.set_ints("mesh.edge_tri", plus1(s.etri), {2, E});
        o.set_reals("mesh.elem_area", s.elem_area, {T});
        o.set_reals("mesh.edge_dxdy", s.edxdy, {2, E});
        o.set_reals("mesh.edge_cross_dxdy", s.ecd, {4, E});
        o.set_reals("mesh.elem_cos", s.elem_cos, {T});
        std::vector<int> nie((size_t)s.max_adj * N, 0);
        for (int n = 0; n < N; ++n)
            for (int k = 0; k < s.nie_num[n]; ++k) nie[(size_t)n * s.max_adj + k] = s.nie[(size_t)n * s.max_adj + k] + 1;
        o.set_ints("mesh.nod_in_elem2D", nie, {s.max_adj, N});
        o.set_ints("mesh.nod_in_elem2D_num", s.nie_num, {N});
        o.set_reals("mesh.gradient_sca", s.gsca, {6, T});
        if (s.have_zbar) o.set_reals("mesh.zbar", s.zbar, {nl});
        o.set_ints("mesh.ulevels", s.ulev_e, {T});
        o.set_ints("mesh.ulevels_nod2D", s.ulev_n, {N});
        o.set_ints("mesh.nlevels", s.nlev_e, {T});
        o.set_ints("mesh.nlevels_nod2D", s.nlev_n, {N});
        // nlevels_nod2D_min: the fewest levels of the node's elements (lib/oce_mesh.F90:1522-1560)
        std::vector<int> nmin(N, nl);
        for (int t = 0; t < T; ++t)
            for (int c = 0; c < 3; ++c) {
                const int n = s.enodes[3 * (size_t)t + c];
                nmin[n] = std::min(nmin[n], s.nlev_e[t]);
            }
        o.set_ints("mesh.nlevels_nod2D_min", nmin, {N});
        std::vector<int> umax(N, 1);  // ulevels_nod2D_max: the deepest top level of its elements
        for (int t = 0; t < T; ++t)
            for (int c = 0; c < 3; ++c) {
                const int n = s.enodes[3 * (size_t)t + c];
                umax[n] = std::max(umax[n], s.ulev_e[t]);
            }
        o.set_ints("mesh.ulevels_nod2D_max", umax, {N});
        o.set_reals("mesh.area", s.area, {nl, N});
        o.set_reals("mesh.areasvol", s.areasvol, {nl, N});
        o.set_reals("mesh.hnode", s.hnode, {L, N});
        o.set_reals("mesh.hnode_new", s.hnode_new, {L, N});
        o.set_reals("mesh.zbar_3d_n", s.zbar3, {nl, N});
        o.set_reals("mesh.Z_3d_n", s.Z, {L, N});
        o.set_reals("mesh.helem", s.helem, {L, T});
        o.set_scalar_int("npes", 1);
        o.set_scalar_int("mype", 0);
        o.set_scalar_int("maxPEnum", 100);
        o.set_scalar_int("myDim_nod2D", N);
        o.set_scalar_int("myDim_elem2D", T);
        o.set_scalar_int("myDim_edge2D", E);
        o.set_scalar_int("dyn.opt_visc", 5);
    }
    o.set_scalar_int("num_tracers", s.ntr);
    auto hor_name = [](int h) { return h == HOR_MUSCL ? "MUSCL" : h == HOR_MFCT ? "MFCT" : "UPW1"; };
    auto ver_name = [](int v) { return v == VER_QR4C ? "QR4C" : v == VER_CDIFF ? "CDIFF" : v == VER_PPM ? "PPM" : "UPW1"; };
    for (int i = 0; i < s.ntr; ++i) {
        const std::string p = "tracer." + std::to_string(i + 1) + ".";
        o.set_reals(p + "values", s.values[i], {L, N});
        // valuesAB: the time level the next step reads (values unless the library keeps it apart)
        const bool ab = s.ab_mode != 0 && i < (int)s.ab_host.size() && !s.ab_host[i].empty();
        o.set_reals(p + "valuesAB", ab ? s.ab_host[i] : s.values[i], {L, N});
        const SchemeSel &sc = s.schemes[i];
        o.set_chars(p + "tra_adv_hor", hor_name(sc.hor));
        o.set_chars(p + "tra_adv_ver", ver_name(sc.ver));
        o.set_chars(p + "tra_adv_lim", sc.fct ? "FCT" : "NONE");
        o.set_scalar_real(p + "tra_adv_ph", sc.ph);
        o.set_scalar_real(p + "tra_adv_pv", sc.pv);
        if (!o.get(p + "ID")) o.set_scalar_int(p + "ID", i + 1);
    }
    if (s.have_muscl) {
        o.set_ints("work.nboundary_lay", s.nbl, {N});
        o.set_ints("work.edge_up_dn_tri", plus1(s.updn), {2, E});
    }
    o.set_logical("dyn.use_wsplit", s.wsplit);
    if (s.have_vel) {
        o.set_reals("dyn.uv", s.uv, {2, L, T});
        o.set_reals("dyn.w", s.w, {nl, N});
        o.set_reals("dyn.w_e", s.we, {nl, N});
        o.set_reals("dyn.w_i", s.wi, {nl, N});
    }
}

// tracer_init / tracer_load_mesh_partition: <dir>/np1 restart files -> the setters.
int load_restart_dir(const char *dir, bool tracers, const char *who) {
    const std::string np = std::string(dir) + "/np1";
    struct stat st;
    if (::stat(np.c_str(), &st) != 0) return fail("could not find restart directory: %s", np.c_str());
    RestartFiles r;
    std::string err;
    if (read_restart(np, tracers, r, err)) return fail("%s: %s", who, err.c_str());
    const int N = r.scalar_int("mesh.nod2D"), E = r.scalar_int("mesh.edge2D"), T = r.scalar_int("mesh.elem2D");
    const int nl = r.scalar_int("mesh.nl"), L = nl - 1;
    auto need = [&](const std::string &name, size_t n) -> bool {
        const RsField *f = r.get(name);
        const size_t have = f ? f->data.size() / (f->kind == 'r' ? 8 : 4) : 0;
        if (have == n) return true;
        fail("%s: %s/%s has %zu entries, expected %zu (one partition, eDim = 0)", who, np.c_str(), name.c_str(), have, n);
        return false;
    };
    const RsField *nief = r.get("mesh.nod_in_elem2D");
    const int max_adj = nief && nief->dims.size() == 2 ? nief->dims[0] : 0;
    if (!need("mesh.elem2D_nodes", 3 * (size_t)T) || !need("mesh.edges", 2 * (size_t)E) ||
        !need("mesh.edge_tri", 2 * (size_t)E) || !need("mesh.nod_in_elem2D", (size_t)max_adj * N) ||
        !need("mesh.nod_in_elem2D_num", N) || !need("mesh.nlevels_nod2D", N) || !need("mesh.ulevels_nod2D", N) ||
        !need("mesh.nlevels", T) || !need("mesh.ulevels", T) || !need("mesh.elem_area", T) ||
        !need("mesh.area", (size_t)nl * N) || !need("mesh.areasvol", (size_t)nl * N) ||
        !need("mesh.edge_dxdy", 2 * (size_t)E) || !need("mesh.edge_cross_dxdy", 4 * (size_t)E) ||
        !need("mesh.elem_cos", T) || !need("mesh.gradient_sca", 6 * (size_t)T) || !need("mesh.hnode", (size_t)L * N) ||
        !need("mesh.hnode_new", (size_t)L * N) || !need("mesh.helem", (size_t)L * T) ||
        !need("mesh.Z_3d_n", (size_t)L * N) || !need("mesh.zbar_3d_n", (size_t)nl * N))
        return -1;
    const auto enodes = r.ints("mesh.elem2D_nodes"), edges = r.ints("mesh.edges"), etri = r.ints("mesh.edge_tri");
    const auto nie = r.ints("mesh.nod_in_elem2D"), nien = r.ints("mesh.nod_in_elem2D_num");
    const auto nlev_n = r.ints("mesh.nlevels_nod2D"), ulev_n = r.ints("mesh.ulevels_nod2D");
    const auto nlev_e = r.ints("mesh.nlevels"), ulev_e = r.ints("mesh.ulevels");
    const auto ea = r.reals("mesh.elem_area"), ar = r.reals("mesh.area"), av = r.reals("mesh.areasvol");
    const auto dxdy = r.reals("mesh.edge_dxdy"), ecd = r.reals("mesh.edge_cross_dxdy"), ec = r.reals("mesh.elem_cos");
    const auto gs = r.reals("mesh.gradient_sca"), hn = r.reals("mesh.hnode"), hnn = r.reals("mesh.hnode_new");
    const auto he = r.reals("mesh.helem"), Z = r.reals("mesh.Z_3d_n"), zb3 = r.reals("mesh.zbar_3d_n"), zb = r.reals("mesh.zbar");
    if (tracer_set_mesh_dims(nl, N, E, T) ||
        tracer_set_partition(r.scalar_int("myDim_nod2D"), r.scalar_int("eDim_nod2D")) ||
        tracer_set_connectivity(enodes.data(), edges.data(), etri.data(), nie.data(), nien.data(), max_adj, N, T, E) ||
        tracer_set_levels(nlev_n.data(), ulev_n.data(), N) || tracer_set_elem_levels(nlev_e.data(), ulev_e.data(), T) ||
        tracer_set_geometry(ea.data(), ar.data(), av.data(), dxdy.data(), ecd.data(), ec.data(), gs.data(), nl, N, T, E) ||
        tracer_set_vertical_grid(hn.data(), hnn.data(), he.data(), Z.data(), zb3.data(), nl, N, T) ||
        (zb.size() == (size_t)nl && tracer_set_zbar(zb.data(), nl)))  // zbar: only tracer_init_ale_arrays reads it
        return -1;
    const auto updn = r.ints("work.edge_up_dn_tri"), nbl = r.ints("work.nboundary_lay");
    if (updn.size() == 2 * (size_t)E && nbl.size() == (size_t)N &&
        tracer_set_muscl_support(updn.data(), nbl.data(), N, E))
        return -1;
    // velocities (not allocated in the file: at rest; w_e / w_i of another shape: w)
    std::vector<double> uv = r.reals("dyn.uv"), w = r.reals("dyn.w"), we = r.reals("dyn.w_e"), wi = r.reals("dyn.w_i");
    if (uv.size() != 2 * (size_t)L * T) uv.assign(2 * (size_t)L * T, 0.0);
    if (w.size() != (size_t)nl * N) w.assign((size_t)nl * N, 0.0);
    if (tracer_set_velocity_elem(uv.data(), w.data(), wi.size() == w.size() ? wi.data() : nullptr,
                                 we.size() == w.size() ? we.data() : nullptr, nl, N, T) ||
        tracer_set_wsplit(r.logical("dyn.use_wsplit") ? 1 : 0))
        return -1;
    if (tracers) {
        const int ntr = r.scalar_int("num_tracers");
        if (ntr < 1) return fail("%s: %s/t_tracer.0 holds no tracers", who, np.c_str());
        if (tracer_allocate_tracers(ntr, nl, N, 2)) return -1;
        for (int i = 1; i <= ntr; ++i) {
            const std::string p = "tracer." + std::to_string(i) + ".";
            if (!need(p + "values", (size_t)L * N)) return -1;
            const auto v = r.reals(p + "values");
            if (tracer_set_values(i, v.data(), nl, N) ||
                tracer_set_scheme(i, r.chars(p + "tra_adv_hor").c_str(), r.chars(p + "tra_adv_ver").c_str(),
                                  r.chars(p + "tra_adv_lim").c_str(), r.scalar_real(p + "tra_adv_ph", 1.0),
                                  r.scalar_real(p + "tra_adv_pv", 1.0)))
                return -1;
        }
    }
    State &s = S();
    if (validate(s)) return -1;
    s.rs = std::move(r);
    s.from_restart = true;
    s.initialized = true;  // the reference's tracer_initialized
    return 0;
}

}  // namespace

// =============================================================================================
extern "C" {

int tracer_init_mpi(int mpi_comm) {
    (void)mpi_comm;
    S();
    return 0;
}

int tracer_finalize_mpi(void) { return 0; }

int tracer_init(const char *restart_dir) {
    if (!restart_dir) return fail("tracer_init: null path");
    return load_restart_dir(restart_dir, true, "tracer_init");
}

int tracer_advect_step(double dt) {
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized. Call tracer_init first.");
    return advect(s, 1, dt);
}

int tracer_advect_steps(int nsteps, double dt) {
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized. Call tracer_init first.");
    if (nsteps < 0) return fail("nsteps must be >= 0");
    return advect(s, nsteps, dt);
}

int tracer_get_stats(double *tmin, double *tmax, double *tsum) {
    State &s = S();
    if (!s.initialized) {
        if (tmin) *tmin = 0.0;
        if (tmax) *tmax = 0.0;
        if (tsum) *tsum = 0.0;
        return fail("tracer not initialized.");
    }
    return stats(s, 1, 1, tmin, tmax, tsum, nullptr);
}

int tracer_get_stats_by_id(int tracer_id, double *tmin, double *tmax, double *tsum, double *vsum) {
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    return stats(s, tracer_id, 0, tmin, tmax, tsum, vsum);
}

int tracer_get_size(int *nz, int *nn) {
    State &s = S();
    if (!s.initialized) {
        if (nz) *nz = 0;
        if (nn) *nn = 0;
        return fail("tracer not initialized.");
    }
    *nz = s.nl - 1;
    *nn = s.N;
    return 0;
}

int tracer_get_values_by_id(int tracer_id, double *values, int nz, int nn) {
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (nz != s.nl - 1 || nn != s.N) return fail("tracer_get_values: expected (%d, %d), got (%d, %d)", s.nl - 1, s.N, nz, nn);
    if (ensure_device(s)) return -1;
    return from_device(s, s.dT[(tracer_id - 1) / s.kb], (tracer_id - 1) % s.kb, values, false);
}

int tracer_get_values(double *values, int nz, int nn) { return tracer_get_values_by_id(1, values, nz, nn); }

int tracer_finalize(const char *output_dir) {
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (!output_dir) return fail("tracer_finalize: null path");
    if (s.nmem > 1)
        return fail("tracer_finalize: the restart files hold one velocity; an ensemble of %d members is not written",
                    s.nmem);
    if (s.on_device && pull_values(s)) return -1;
    RestartFiles out;
    restart_of_state(s, out);
    std::string err;
    const std::string dir(output_dir);
    if (write_restart(dir + "/np1", dir + "/meta.time", out, s.ntr, err)) return fail("tracer_finalize: %s", err.c_str());
    std::printf("Restart files written to: %s/np1\n", output_dir);
    std::fflush(stdout);
    s.initialized = false;  // as the reference: initialise again before stepping
    return 0;
}

int tracer_run_workflow(const char *restart_dir, int nsteps, double dt) {
    if (tracer_init(restart_dir)) return -1;
    for (int step = 1; step <= nsteps; ++step) {
        double tmin, tmax, tsum;
        if (tracer_advect_step(dt) || tracer_get_stats(&tmin, &tmax, &tsum)) return -1;
        std::printf("Step %3d: min/max/sum = %15.6E%15.6E%15.6E\n", step, tmin, tmax, tsum);
    }
    std::fflush(stdout);
    return tracer_finalize(restart_dir);
}

int tracer_cleanup(void) {
    if (g) {
        release_device(*g);
        if (g->ev0) (void)hipEventDestroy(g->ev0);
        if (g->ev1) (void)hipEventDestroy(g->ev1);
        if (g->cfg.fork) (void)hipEventDestroy(g->cfg.fork);
        if (g->cfg.join) (void)hipEventDestroy(g->cfg.join);
        if (g->cfg.aux) (void)hipStreamDestroy(g->cfg.aux);
        if (g->stream) (void)hipStreamDestroy(g->stream);
        delete g;
        g = nullptr;
    }
    g_err.clear();
    return 0;
}

int tracer_set_mesh_dims(int nl, int nod2D, int edge2D, int elem2D) {
    if (nl < 3 || nod2D < 1 || edge2D < 1 || elem2D < 1)
        return fail("tracer_set_mesh_dims: invalid dims nl=%d nod2D=%d edge2D=%d elem2D=%d", nl, nod2D, edge2D, elem2D);
    State &s = S();
    release_device(s);
    const hipStream_t keep = s.stream;
    hipEvent_t k0 = s.ev0, k1 = s.ev1;
    const int prec = s.prec, kbr = s.kb_req;
    const bool graph = s.use_graph, sr = s.stochastic, fused = s.fused, ht = s.halo_tiles, ws = s.wsplit;
    const bool comp = s.compensated, flim = s.fused_limiter, keepw = s.keep_work, conc = s.concurrent;
    const int abm = s.ab_mode;
    const double abe = s.ab_eps;
    const uint64_t gen = s.gen;
    const LaunchCfg cfg = s.cfg;
    s = State();
    s.cfg = cfg;
    s.stream = keep;
    s.ev0 = k0;
    s.ev1 = k1;
    s.prec = prec;
    s.kb_req = kbr;
    s.use_graph = graph;
    s.stochastic = sr;
    s.compensated = comp;
    s.fused_limiter = flim;
    s.keep_work = keepw;
    s.concurrent = conc;
    s.fused = fused;
    s.halo_tiles = ht;
    s.wsplit = ws;
    s.ab_mode = abm;
    s.ab_eps = abe;
    s.gen = gen + 1;
    s.nl = nl;
    s.N = nod2D;
    s.E = edge2D;
    s.T = elem2D;
    s.dims = true;
    return 0;
}

int tracer_set_levels(const int *nlevels, const int *ulevels, int nod2D) {
    State &s = S();
    if (!s.dims) return fail("tracer_set_levels: call tracer_set_mesh_dims first");
    if (nod2D != s.N) return fail("tracer_set_levels: nod2D=%d != %d", nod2D, s.N);
    s.nlev_n.assign(nlevels, nlevels + nod2D);
    s.ulev_n.assign(ulevels, ulevels + nod2D);
    s.have_nlev = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_elem_levels(const int *nlevels, const int *ulevels, int elem2D) {
    State &s = S();
    if (!s.dims) return fail("tracer_set_elem_levels: call tracer_set_mesh_dims first");
    if (elem2D != s.T) return fail("tracer_set_elem_levels: elem2D=%d != %d", elem2D, s.T);
    s.nlev_e.assign(nlevels, nlevels + elem2D);
    s.ulev_e.assign(ulevels, ulevels + elem2D);
    s.have_elev = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_thickness(const double *hnode, int nl, int nod2D) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_set_thickness")) return -1;
    s.hnode.assign(hnode, hnode + (size_t)(nl - 1) * nod2D);
    s.hnode_new = s.hnode;
    s.have_hnode = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_partition(int myDim_nod2D, int eDim_nod2D) {
    State &s = S();
    if (eDim_nod2D != 0) return fail("tracer_set_partition: eDim_nod2D=%d; one partition without halo is supported", eDim_nod2D);
    if (s.dims && myDim_nod2D != s.N) return fail("tracer_set_partition: myDim_nod2D=%d != nod2D=%d", myDim_nod2D, s.N);
    s.myDim = myDim_nod2D;
    return 0;
}

int tracer_allocate_tracers(int num_tracers, int nl, int nod2D, int AB_order) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_allocate_tracers")) return -1;
    if (num_tracers < 1) return fail("tracer_allocate_tracers: num_tracers=%d", num_tracers);
    if (AB_order < 1) return fail("tracer_allocate_tracers: AB_order=%d", AB_order);
    s.ntr = num_tracers;
    s.values.assign(num_tracers, std::vector<double>((size_t)(nl - 1) * nod2D, 0.0));
    s.schemes.assign(num_tracers, SchemeSel());
    s.tr_mem.assign(num_tracers, 0);
    s.on_device = false;
    return 0;
}

int tracer_set_values(int tracer_id, const double *values, int nl, int nod2D) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (check_dims(s, nl, nod2D, "tracer_set_values")) return -1;
    std::vector<double> &v = s.values[tracer_id - 1];
    v.assign(values, values + (size_t)(nl - 1) * nod2D);
    // values, and valuesAB / valuesold where held apart (lib/tracer_c_interface.F90:613-648)
    if (s.on_device) {
        const int b = (tracer_id - 1) / s.kb, lane = (tracer_id - 1) % s.kb;
        if (to_device(s, v.data(), false, s.dT[b], lane)) return -1;
        if (s.tab_on[b] && to_device(s, v.data(), false, s.tabbuf[b], lane)) return -1;
        if (s.ab_mode == 2 && s.toldbuf[b] && to_device(s, v.data(), false, s.toldbuf[b], lane)) return -1;
        s.ab_fresh = true;
        touch_values(s);
    } else {
        if (tracer_id - 1 < (int)s.ab_host.size()) s.ab_host[tracer_id - 1].clear();
        if (tracer_id - 1 < (int)s.old_host.size()) s.old_host[tracer_id - 1].clear();
    }
    return 0;
}

int tracer_set_values_ab(int tracer_id, const double *values, int nl, int nod2D) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (check_dims(s, nl, nod2D, "tracer_set_values_ab")) return -1;
    if (ensure_device(s)) return -1;
    const int b = (tracer_id - 1) / s.kb, lane = (tracer_id - 1) % s.kb;
    const size_t bytes = (size_t)(nl - 1) * s.Np * s.kb * storage_bytes(s.prec);
    if (!s.tabbuf[b] && dalloc_bytes(s, &s.tabbuf[b], bytes)) return -1;
    if (!s.tab_on[b]) {  // the block's other lanes keep valuesAB = values
        HIPCHK(hipMemcpyAsync(s.tabbuf[b], s.dT[b], bytes, hipMemcpyDeviceToDevice, s.stream));
        s.tab_on[b] = 1;
    }
    return to_device(s, values, false, s.tabbuf[b], lane);
}

int tracer_set_velocity_zero(int nl, int nod2D, int edge2D) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_set_velocity_zero")) return -1;
    // edge2D only sizes the reference's own (unused) edge array; every velocity here is zeroed
    if (edge2D < 0) return fail("tracer_set_velocity_zero: edge2D=%d", edge2D);
    s.uv.assign((size_t)2 * (nl - 1) * s.T, 0.0);
    s.w.assign((size_t)nl * nod2D, 0.0);
    s.wi = s.w;
    s.we = s.w;
    s.have_vel = true;
    return refresh_velocity(s);
}

int tracer_set_velocity(const double *uv, const double *w, int nl, int nod2D, int edge2D) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_set_velocity")) return -1;
    for (size_t i = 0; i < (size_t)(nl - 1) * edge2D * 2; ++i)
        if (uv[i] != 0.0)
            return fail("tracer_set_velocity: uv(nl-1, edge2D, 2) is not the element velocity "
                        "vel(2, nl-1, elem2D) that do_oce_adv_tra reads; use tracer_set_velocity_elem");
    s.uv.assign((size_t)2 * (nl - 1) * s.T, 0.0);
    s.w.assign(w, w + (size_t)nl * nod2D);
    s.wi = s.w;
    s.we = s.w;
    s.have_vel = true;
    return refresh_velocity(s);
}

int tracer_set_velocity_elem(const double *uv, const double *w, const double *w_i, const double *w_e,
                             int nl, int nod2D, int elem2D) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_set_velocity_elem")) return -1;
    if (elem2D != s.T) return fail("tracer_set_velocity_elem: elem2D=%d != %d", elem2D, s.T);
    s.uv.assign(uv, uv + (size_t)2 * (nl - 1) * elem2D);
    s.w.assign(w, w + (size_t)nl * nod2D);
    s.wi = w_i ? std::vector<double>(w_i, w_i + (size_t)nl * nod2D) : s.w;
    s.we = w_e ? std::vector<double>(w_e, w_e + (size_t)nl * nod2D) : s.w;
    s.have_vel = true;
    return refresh_velocity(s);
}

int tracer_init_complete(void) {
    State &s = S();
    if (!s.have_vel && s.dims) {
        s.uv.assign((size_t)2 * (s.nl - 1) * s.T, 0.0);
        s.w.assign((size_t)s.nl * s.N, 0.0);
        s.wi = s.w;
        s.we = s.w;
        s.have_vel = true;
    }
    if (validate(s)) return -1;
    s.initialized = true;
    return ensure_device(s);
}

int tracer_load_mesh_partition(const char *restart_dir) {
    if (!restart_dir) return fail("tracer_load_mesh_partition: null path");
    return load_restart_dir(restart_dir, false, "tracer_load_mesh_partition");
}

int tracer_load_mesh_from_files(const char *mesh_path) {
    if (!mesh_path) return fail("tracer_load_mesh_from_files: null path");
    FileMesh fm;
    std::string err;
    if (read_mesh_dir(mesh_path, fm, err)) return fail("tracer_load_mesh_from_files(%s): %s", mesh_path, err.c_str());
    // hand the arrays to the setters (their checks apply); one partition holds the whole mesh
    if (tracer_set_mesh_dims(fm.nl, fm.N, fm.E, fm.T) || tracer_set_partition(fm.N, 0) ||
        tracer_set_coordinates(fm.coord.data(), fm.N) ||
        tracer_set_connectivity(fm.enodes.data(), fm.edges.data(), fm.edge_tri.data(), fm.nie.data(),
                                fm.nie_num.data(), fm.max_adj, fm.N, fm.T, fm.E) ||
        tracer_set_levels(fm.nlev_n.data(), fm.ulev_n.data(), fm.N) ||
        tracer_set_elem_levels(fm.nlev_e.data(), fm.ulev_e.data(), fm.T) ||
        tracer_set_geometry(fm.elem_area.data(), fm.area.data(), fm.areasvol.data(), fm.edxdy.data(), fm.ecd.data(),
                            fm.elem_cos.data(), fm.gsca.data(), fm.nl, fm.N, fm.T, fm.E) ||
        tracer_set_vertical_grid(fm.hnode.data(), fm.hnode.data(), fm.helem.data(), fm.Z.data(), fm.zbar3.data(), fm.nl,
                                 fm.N, fm.T) ||
        tracer_set_zbar(fm.zbar.data(), fm.nl) || tracer_set_muscl_support(fm.updn.data(), fm.nbl.data(), fm.N, fm.E))
        return -1;
    // as the reference (:830-870), a loaded mesh counts as initialised: tracer_get_size, the ALE
    // and advection set-up calls and stepping are available before tracer_init_complete
    if (validate(S())) return -1;
    S().initialized = true;
    return 0;
}

int tracer_set_zbar(const double *zbar, int nl) {
    State &s = S();
    if (!s.dims || nl != s.nl) return fail("tracer_set_zbar: nl=%d does not match the mesh", nl);
    s.zbar.assign(zbar, zbar + nl);
    s.have_zbar = true;
    return 0;
}

int tracer_init_ale_arrays(void) {
    State &s = S();
    if (!s.dims) return fail("tracer_init_ale_arrays called before initialization");
    if (!s.have_zbar) return fail("tracer_init_ale_arrays: zbar not set (tracer_set_zbar)");
    const int L = s.nl - 1;
    s.hnode.assign((size_t)L * s.N, 0.0);
    s.helem.assign((size_t)L * s.T, 0.0);
    s.Z.assign((size_t)L * s.N, 0.0);
    s.zbar3.assign((size_t)s.nl * s.N, 0.0);
    for (int n = 0; n < s.N; ++n) {
        for (int k = 0; k < L; ++k) {
            s.hnode[(size_t)n * L + k] = s.zbar[k] - s.zbar[k + 1];
            s.Z[(size_t)n * L + k] = 0.5 * (s.zbar[k] + s.zbar[k + 1]);
        }
        for (int k = 0; k < s.nl; ++k) s.zbar3[(size_t)n * s.nl + k] = s.zbar[k];
    }
    for (int t = 0; t < s.T; ++t)
        for (int k = 0; k < L; ++k) s.helem[(size_t)t * L + k] = s.zbar[k] - s.zbar[k + 1];
    s.hnode_new = s.hnode;
    s.have_hnode = true;
    s.have_vert = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_init_for_advection(void) {
    State &s = S();
    if (!s.initialized) return fail("tracer_init_for_advection called before mesh loaded");
    return ensure_device(s);
}

int tracer_set_coordinates(const double *coord_nod2D, int nod2D) {
    State &s = S();
    if (!s.dims || nod2D != s.N) return fail("tracer_set_coordinates: nod2D=%d does not match", nod2D);
    s.coord.assign(coord_nod2D, coord_nod2D + 2 * (size_t)nod2D);
    s.have_coord = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_connectivity(const int *elem2D_nodes, const int *edges, const int *edge_tri,
                            const int *nod_in_elem2D, const int *nod_in_elem2D_num, int max_adj,
                            int nod2D, int elem2D, int edge2D) {
    State &s = S();
    if (!s.dims) return fail("tracer_set_connectivity: call tracer_set_mesh_dims first");
    if (nod2D != s.N || elem2D != s.T || edge2D != s.E)
        return fail("tracer_set_connectivity: dims (%d, %d, %d) do not match the mesh", nod2D, elem2D, edge2D);
    if (max_adj < 1) return fail("tracer_set_connectivity: max_adj=%d", max_adj);
    s.max_adj = max_adj;
    s.enodes.resize((size_t)3 * elem2D);
    for (size_t i = 0; i < s.enodes.size(); ++i) {
        const int v = elem2D_nodes[i];
        if (v < 1 || v > nod2D) return fail("elem2D_nodes(%zu, %zu)=%d out of range", i % 3 + 1, i / 3 + 1, v);
        s.enodes[i] = v - 1;
    }
    s.edges.resize((size_t)2 * edge2D);
    s.etri.resize((size_t)2 * edge2D);
    for (size_t i = 0; i < s.edges.size(); ++i) {
        const int v = edges[i], t = edge_tri[i];
        if (v < 1 || v > nod2D) return fail("edges(%zu, %zu)=%d out of range", i % 2 + 1, i / 2 + 1, v);
        if (t < 0 || t > elem2D) return fail("edge_tri(%zu, %zu)=%d out of range", i % 2 + 1, i / 2 + 1, t);
        s.edges[i] = v - 1;
        s.etri[i] = t - 1;
    }
    s.nie.assign((size_t)max_adj * nod2D, 0);
    s.nie_num.assign(nod_in_elem2D_num, nod_in_elem2D_num + nod2D);
    for (int n = 0; n < nod2D; ++n) {
        if (s.nie_num[n] < 1 || s.nie_num[n] > max_adj)
            return fail("nod_in_elem2D_num(%d)=%d outside [1, %d]", n + 1, s.nie_num[n], max_adj);
        for (int k = 0; k < s.nie_num[n]; ++k) {
            const int t = nod_in_elem2D[(size_t)n * max_adj + k];
            if (t < 1 || t > elem2D) return fail("nod_in_elem2D(%d, %d)=%d out of range", k + 1, n + 1, t);
            s.nie[(size_t)n * max_adj + k] = t - 1;
        }
    }
    s.have_conn = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_geometry(const double *elem_area, const double *area, const double *areasvol,
                        const double *edge_dxdy, const double *edge_cross_dxdy,
                        const double *elem_cos, const double *gradient_sca, int nl, int nod2D,
                        int elem2D, int edge2D) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_set_geometry")) return -1;
    if (elem2D != s.T || edge2D != s.E) return fail("tracer_set_geometry: elem2D/edge2D do not match");
    s.elem_area.assign(elem_area, elem_area + elem2D);
    s.area.assign(area, area + (size_t)nl * nod2D);
    s.areasvol.assign(areasvol, areasvol + (size_t)nl * nod2D);
    s.edxdy.assign(edge_dxdy, edge_dxdy + (size_t)2 * edge2D);
    s.ecd.assign(edge_cross_dxdy, edge_cross_dxdy + (size_t)4 * edge2D);
    s.elem_cos.assign(elem_cos, elem_cos + elem2D);
    s.gsca.assign(gradient_sca, gradient_sca + (size_t)6 * elem2D);
    s.have_geom = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_vertical_grid(const double *hnode, const double *hnode_new, const double *helem,
                             const double *Z_3d_n, const double *zbar_3d_n, int nl, int nod2D,
                             int elem2D) {
    State &s = S();
    if (check_dims(s, nl, nod2D, "tracer_set_vertical_grid")) return -1;
    if (elem2D != s.T) return fail("tracer_set_vertical_grid: elem2D=%d != %d", elem2D, s.T);
    const size_t NL = (size_t)(nl - 1) * nod2D;
    s.hnode.assign(hnode, hnode + NL);
    s.hnode_new.assign(hnode_new ? hnode_new : hnode, (hnode_new ? hnode_new : hnode) + NL);
    s.helem.assign(helem, helem + (size_t)(nl - 1) * elem2D);
    s.Z.assign(Z_3d_n, Z_3d_n + NL);
    s.zbar3.assign(zbar_3d_n, zbar_3d_n + (size_t)nl * nod2D);
    s.have_hnode = true;
    s.have_vert = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_muscl_support(const int *edge_up_dn_tri, const int *nboundary_lay, int nod2D, int edge2D) {
    State &s = S();
    if (!s.dims || nod2D != s.N || edge2D != s.E) return fail("tracer_set_muscl_support: dims do not match");
    s.updn.resize((size_t)2 * edge2D);
    for (size_t i = 0; i < s.updn.size(); ++i) {
        const int t = edge_up_dn_tri[i];
        if (t < 0 || t > s.T) return fail("edge_up_dn_tri(%zu, %zu)=%d out of range", i % 2 + 1, i / 2 + 1, t);
        s.updn[i] = t - 1;
    }
    s.nbl.assign(nboundary_lay, nboundary_lay + nod2D);
    s.have_muscl = true;
    if (invalidate(s)) return -1;
    return 0;
}

int tracer_set_scheme(int tracer_id, const char *hor, const char *ver, const char *lim, double ph, double pv) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    SchemeSel sc;
    if (!hor || !parse_hor(hor, sc.hor)) return fail("Unknown horizontal advection type %s! Check your namelists!", hor ? hor : "(null)");
    if (!ver || !parse_ver(ver, sc.ver)) return fail("Unknown vertical advection type %s! Check your namelists!", ver ? ver : "(null)");
    if (!lim || (strcmp(lim, "FCT") && strcmp(lim, "NONE"))) return fail("Unknown limiter %s", lim ? lim : "(null)");
    sc.fct = !strcmp(lim, "FCT");
    sc.ph = ph;
    sc.pv = pv;
    if ((sc.hor == HOR_MUSCL || sc.hor == HOR_MFCT) && !s.have_muscl)
        return fail("tracer %d: %s needs edge_up_dn_tri (tracer_set_muscl_support)", tracer_id, hor);
    s.schemes[tracer_id - 1] = sc;
    s.gen++;
    return 0;
}

int tracer_set_precision(int mode) {
    State &s = S();
    if (mode < PREC_F64 || mode > PREC_F32_BF16) return fail("tracer_set_precision: unknown mode %d", mode);
    if (s.prec == mode) return 0;
    if (s.on_device) {
        // keep the current values: read them back at the old precision, rebuild at the new one
        if (pull_values(s)) return -1;
        release_device(s);
    }
    s.prec = mode;
    s.gen++;
    if (s.initialized) return ensure_device(s);
    return 0;
}

int tracer_set_tracers_per_pass(int kb) {
    State &s = S();
    if (kb != 0 && (kb < 1 || kb > kMaxKB || (kb & (kb - 1))))
        return fail("tracer_set_tracers_per_pass: %d is not 0 (automatic) or a power of two in [1, %d]", kb, kMaxKB);
    const int old = s.kb_req;
    s.kb_req = kb;
    if (s.on_device && block_width(s) != s.kb) {  // re-interleave the stored values
        const int want = s.kb_req;
        s.kb_req = old;
        if (pull_values(s)) return -1;
        s.kb_req = want;
        release_device(s);
        if (ensure_device(s)) return -1;
    }
    s.gen++;
    return 0;
}

int tracer_get_tracers_per_pass(int *kb) {
    const State &s = S();
    *kb = s.on_device ? s.kb : block_width(s);
    return 0;
}

int tracer_checkpoint(int restore) {
    State &s = S();
    if (ensure_device(s)) return -1;
    const size_t bytes = (size_t)(s.nl - 1) * s.Np * s.kb * storage_bytes(s.prec);
    const size_t rbytes = (size_t)(s.nl - 1) * s.Np * s.kb * sizeof(float);  // FP32 remainders
    if (!restore) {
        if (s.ckpt.size() != s.dT.size()) s.ckpt.assign(s.dT.size(), nullptr);
        if (s.ckpt_res.size() != s.dT.size()) s.ckpt_res.assign(s.dT.size(), nullptr);
        if (s.ckpt_ab.size() != s.dT.size()) s.ckpt_ab.assign(s.dT.size(), nullptr);
        if (s.ckpt_old.size() != s.dT.size()) s.ckpt_old.assign(s.dT.size(), nullptr);
        for (size_t b = 0; b < s.dT.size(); ++b) {
            if (!s.ckpt[b] && dalloc_bytes(s, &s.ckpt[b], bytes)) return -1;
            HIPCHK(hipMemcpyAsync(s.ckpt[b], s.dT[b], bytes, hipMemcpyDeviceToDevice, s.stream));
            if (b < s.dR.size() && s.dR[b]) {
                if (!s.ckpt_res[b] && dalloc_bytes(s, &s.ckpt_res[b], rbytes)) return -1;
                HIPCHK(hipMemcpyAsync(s.ckpt_res[b], s.dR[b], rbytes, hipMemcpyDeviceToDevice, s.stream));
            }
            if (s.ab_mode != 0) {  // the time levels the next steps read
                if (!s.ckpt_ab[b] && dalloc_bytes(s, &s.ckpt_ab[b], bytes)) return -1;
                HIPCHK(hipMemcpyAsync(s.ckpt_ab[b], s.tabbuf[b], bytes, hipMemcpyDeviceToDevice, s.stream));
                if (s.ab_mode == 2) {
                    if (!s.ckpt_old[b] && dalloc_bytes(s, &s.ckpt_old[b], bytes)) return -1;
                    HIPCHK(hipMemcpyAsync(s.ckpt_old[b], s.toldbuf[b], bytes, hipMemcpyDeviceToDevice, s.stream));
                }
            }
        }
        s.ckpt_seed = s.step_seed;
        s.ckpt_ab_mode = s.ab_mode;
    } else {
        if (s.ckpt.size() != s.dT.size()) return fail("tracer_checkpoint: no checkpoint of the current storage");
        if (s.ckpt_ab_mode != s.ab_mode) return fail("tracer_checkpoint: the valuesAB mode changed since the checkpoint");
        for (size_t b = 0; b < s.dT.size(); ++b) {
            HIPCHK(hipMemcpyAsync(s.dT[b], s.ckpt[b], bytes, hipMemcpyDeviceToDevice, s.stream));
            if (b < s.dR.size() && s.dR[b]) {
                if (b >= s.ckpt_res.size() || !s.ckpt_res[b])
                    return fail("tracer_checkpoint: the checkpoint holds no remainders (compensated storage changed)");
                HIPCHK(hipMemcpyAsync(s.dR[b], s.ckpt_res[b], rbytes, hipMemcpyDeviceToDevice, s.stream));
            }
            if (s.ab_mode != 0)
                HIPCHK(hipMemcpyAsync(s.tabbuf[b], s.ckpt_ab[b], bytes, hipMemcpyDeviceToDevice, s.stream));
            if (s.ab_mode == 2)
                HIPCHK(hipMemcpyAsync(s.toldbuf[b], s.ckpt_old[b], bytes, hipMemcpyDeviceToDevice, s.stream));
        }
        s.ab_fresh = true;  // recomputing valuesAB from the restored levels is exact
        if (s.ab_mode == 0) std::fill(s.tab_on.begin(), s.tab_on.end(), 0);
        touch_values(s);
        s.step_seed = s.ckpt_seed;  // the restored state continues the same rounding sequence
    }
    HIPCHK(hipStreamSynchronize(s.stream));
    return 0;
}

int tracer_set_ab_mode(int mode, double epsilon) {
    State &s = S();
    if (mode < 0 || mode > 2) return fail("tracer_set_ab_mode: unknown mode %d", mode);
    if (!(epsilon >= 0.0 && epsilon <= 1.0)) return fail("tracer_set_ab_mode: epsilon=%g outside [0, 1]", epsilon);
    if (mode == s.ab_mode && epsilon == s.ab_eps) return 0;
    s.ab_mode = mode;
    s.ab_eps = epsilon;
    s.ab_host.clear();
    s.old_host.clear();
    s.gen++;
    if (!s.on_device) return 0;
    if (mode == 0) {
        std::fill(s.tab_on.begin(), s.tab_on.end(), 0);
        return 0;
    }
    return ab_prepare(s);  // valuesAB (and valuesold) := values
}

int tracer_get_ab_mode(int *mode, double *epsilon) {
    const State &s = S();
    if (mode) *mode = s.ab_mode;
    if (epsilon) *epsilon = s.ab_eps;
    return 0;
}

int tracer_set_wsplit(int enable) {
    State &s = S();
    s.wsplit = enable != 0;
    s.gen++;
    return 0;
}

int tracer_get_wsplit(int *enabled) {
    if (!enabled) return fail("tracer_get_wsplit: null pointer");
    *enabled = S().wsplit ? 1 : 0;
    return 0;
}

int tracer_set_fused(int enable) {
    State &s = S();
    s.fused = enable != 0;
    s.gen++;
    if (s.fused && s.initialized) {
        if (ensure_device(s)) return -1;
        if (!s.tiles_built && build_tiles(s)) {
            s.fused = false;
            return -1;
        }
    }
    return 0;
}

int tracer_set_halo_tiles(int enable) {
    State &s = S();
    s.halo_tiles = enable != 0;
    s.m64.ht_n = s.halo_tiles ? s.ht_built : 0;
    s.m64.et_n = s.halo_tiles ? s.et_built : 0;
    s.m64.rt_n = (s.halo_tiles && s.fused_limiter) ? s.rt_built : 0;
    s.gen++;
    return 0;
}

int tracer_get_halo_tiles(int *ntiles) {
    if (!ntiles) return fail("tracer_get_halo_tiles: null pointer");
    *ntiles = S().m64.ht_n;
    return 0;
}

int tracer_set_compensated(int enable) {
    State &s = S();
    const bool on = enable != 0;
    if (s.compensated == on) return 0;
    if (s.on_device) {  // keep the values: read back (with remainders), rebuild with or without them
        if (pull_values(s)) return -1;
        release_device(s);
    }
    s.compensated = on;
    s.gen++;
    if (s.initialized) return ensure_device(s);
    return 0;
}

int tracer_set_concurrent(int enable) {
    State &s = S();
    s.concurrent = enable != 0;
    if (s.concurrent && s.stream && !s.cfg.aux) {
        HIPCHK(hipStreamCreateWithFlags(&s.cfg.aux, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&s.cfg.fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&s.cfg.join, hipEventDisableTiming));
    }
    s.gen++;
    return 0;
}

int tracer_set_fused_limiter(int enable) {
    State &s = S();
    s.fused_limiter = enable != 0;
    s.m64.rt_n = (s.halo_tiles && s.fused_limiter) ? s.rt_built : 0;
    s.gen++;
    return 0;
}

int tracer_get_fused_limiter(int *active) {
    if (!active) return fail("tracer_get_fused_limiter: null pointer");
    State &s = S();
    *active = (s.initialized && s.on_device && limiter_applies(s, launch_plan(s))) ? 1 : 0;
    return 0;
}

int tracer_set_keep_work(int keep) {
    S().keep_work = keep != 0;
    return 0;
}

int tracer_set_rounding(int stochastic) {
    State &s = S();
    s.stochastic = stochastic != 0;
    s.gen++;
    return 0;
}

int tracer_set_graph(int enable) {
    State &s = S();
    s.use_graph = enable != 0;
    if (!s.use_graph) drop_graph(s);
    return 0;
}

int tracer_set_launch_config(int lc_edge, int lc_lo, int lc_bounds, int lc_final) {
    if (lc_edge < 1 || lc_lo < 1 || lc_bounds < 1 || lc_final < 1)
        return fail("tracer_set_launch_config: levels per thread must be >= 1");
    State &s = S();
    s.cfg.lc_edge = lc_edge;
    s.cfg.lc_lo = lc_lo;
    s.cfg.lc_bounds = lc_bounds;
    s.cfg.lc_final = lc_final;
    s.gen++;
    return 0;
}

int tracer_get_launch_config(int *cfg4) {
    const State &s = S();
    cfg4[0] = s.cfg.lc_edge;
    cfg4[1] = s.cfg.lc_lo;
    cfg4[2] = s.cfg.lc_bounds;
    cfg4[3] = s.cfg.lc_final;
    return 0;
}

int tracer_zero_tendencies(void) {
    State &s = S();
    if (ensure_device(s) || ensure_dropin_buffers(s)) return -1;
    // no memset: the next drop-in call starts its sums from zero (acc_zero), and a read-back
    // before it zeroes the arrays then
    s.acc_zero = true;
    return 0;
}

// The accumulators' contents for a read-back: zeroed now if tracer_zero_tendencies came last.
static int settle_tendencies(State &s) {
    if (!s.acc_zero) return 0;
    const size_t NLn = (size_t)(s.nl - 1) * s.Np;
    HIPCHK(hipMemsetAsync(s.acc_h, 0, NLn * sizeof(double), s.stream));
    HIPCHK(hipMemsetAsync(s.acc_v, 0, NLn * sizeof(double), s.stream));
    s.acc_zero = false;
    return 0;
}

// The reference zeroes its flux and low-order work arrays at every call; here they are zeroed
// once when allocated and the kernels write exactly the levels each edge / node has (the same set
// at every call), so the levels below stay zero.  The edge transport depends on the velocity
// only: it is computed once per velocity (tracer_set_velocity_elem bumps the version), not once
// per tracer.  The kernels run on the library's stream without a host synchronisation: every
// call that reads results (tracer_get_tendencies, tracer_get_work, ...) waits for them.
int tracer_do_oce_adv_tra(double dt, int tracer_id) {
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (ensure_device(s) || ensure_dropin_buffers(s) || ensure_wsplit(s)) return -1;
    if (s.lo_dirty) {  // a snapshot of values was kept in the LO buffer, also below the bottom
        HIPCHK(hipMemsetAsync(s.LO, 0, (size_t)(s.nl - 1) * s.Np * s.kb * compute_bytes(s.prec), s.stream));
        s.lo_dirty = false;
    }
    const int id = tracer_id - 1, lane = id % s.kb;
    hipError_t e = hipSuccess;
    if (s.q_vel != s.vel_gen) {
        e = launch_qflux(s.m64, s.stream);
        s.q_vel = s.vel_gen;
    }
    if (e == hipSuccess) e = launch_block(s, id / s.kb, 1ull << lane, s.schemes[id], dt, true, nullptr);
    if (e != hipSuccess) {
        s.q_vel = ~0ull;
        return fail("kernel launch failed: %s", hipGetErrorString(e));
    }
    s.last_fct = s.schemes[id].fct;
    s.last_lane = lane;
    s.dropin_valid = true;
    s.work_kept = s.keep_work;
    s.acc_zero = false;  // this call's update wrote (or added to) every wet cell
    return 0;
}

int tracer_get_tendencies(double *dttf_h, double *dttf_v, int nz, int nn) {
    State &s = S();
    if (!s.on_device || !s.acc_h) return fail("tracer_get_tendencies: no tendencies computed yet");
    if (nz != s.nl - 1 || nn != s.N) return fail("tracer_get_tendencies: expected (%d, %d)", s.nl - 1, s.N);
    if (settle_tendencies(s)) return -1;
    int rc = 0;  // accumulators are double, one tracer
    if (dttf_h) rc |= download_work(s, s.acc_h, nz, s.Np, s.perm_n, dttf_h, 8, 1, 0);
    if (dttf_v) rc |= download_work(s, s.acc_v, nz, s.Np, s.perm_n, dttf_v, 8, 1, 0);
    return rc ? -1 : 0;
}

int tracer_get_work(const char *name, double *out, size_t count) {
    State &s = S();
    if (!s.on_device || !s.dropin_valid)
        return fail("tracer_get_work: no tracer_do_oce_adv_tra call since the device state was built or last "
                    "advanced (tracer_advect_step(s) reuse the work arrays for every tracer)");
    const size_t L = s.nl - 1, cb = compute_bytes(s.prec);
    const int kb = s.kb, lane = s.last_lane;
    if (!strcmp(name, "fct_LO")) {
        if (count != L * s.N) return fail("tracer_get_work(fct_LO): count must be %zu", L * s.N);
        return download_work(s, s.LO, (int)L, s.Np, s.perm_n, out, cb, kb, lane);
    }
    if (s.last_fct && !s.work_kept && (!strcmp(name, "adv_flux_hor") || !strcmp(name, "adv_flux_ver")))
        return fail("tracer_get_work(%s): the last tracer_do_oce_adv_tra call did not keep the limited fluxes "
                    "(tracer_set_keep_work(0))", name);
    if (!strcmp(name, "adv_flux_hor")) {
        if (count != L * s.E) return fail("tracer_get_work(adv_flux_hor): count must be %zu", L * s.E);
        if (s.last_fct && s.dbg_h) return download_work(s, s.dbg_h, (int)L, s.Ep, s.perm_e, out, cb, 1, 0);
        if (s.prec != PREC_F64)  // the single-precision kernels store the high-order flux minus the upwind flux
            return fail("tracer_get_work(adv_flux_hor): without the limiter the single-precision modes do not "
                        "store the high-order flux itself");
        return download_work(s, s.adfh, (int)L, s.Ep, s.perm_e, out, cb, kb, lane);
    }
    if (!strcmp(name, "adv_flux_ver")) {
        if (count != (size_t)s.nl * s.N) return fail("tracer_get_work(adv_flux_ver): count must be %zu", (size_t)s.nl * s.N);
        if (s.last_fct && s.dbg_v) return download_work(s, s.dbg_v, s.nl, s.Np, s.perm_n, out, cb, 1, 0);
        return download_work(s, s.adfv, s.nl, s.Np, s.perm_n, out, cb, kb, lane);
    }
    return fail("tracer_get_work: unknown array %s", name);
}

int tracer_set_values_device(int tracer_id, const void *values, int nl, int nod2D) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (check_dims(s, nl, nod2D, "tracer_set_values_device")) return -1;
    if (ensure_device(s)) return -1;
    const int b = (tracer_id - 1) / s.kb, lane = (tracer_id - 1) % s.kb;
    if (to_device(s, static_cast<const double *>(values), true, s.dT[b], lane)) return -1;
    if (s.tab_on[b] && to_device(s, static_cast<const double *>(values), true, s.tabbuf[b], lane)) return -1;
    if (s.ab_mode == 2 && s.toldbuf[b] && to_device(s, static_cast<const double *>(values), true, s.toldbuf[b], lane))
        return -1;
    s.ab_fresh = true;
    touch_values(s);
    return 0;
}

int tracer_get_values_device(int tracer_id, void *values, int nz, int nn) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (nz != s.nl - 1 || nn != s.N) return fail("tracer_get_values_device: expected (%d, %d)", s.nl - 1, s.N);
    if (ensure_device(s)) return -1;
    return from_device(s, s.dT[(tracer_id - 1) / s.kb], (tracer_id - 1) % s.kb, static_cast<double *>(values), true);
}

int tracer_set_values_ab_device(int tracer_id, const void *values, int nl, int nod2D) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (check_dims(s, nl, nod2D, "tracer_set_values_ab_device")) return -1;
    if (!values) return fail("tracer_set_values_ab_device: null pointer");
    if (ensure_device(s)) return -1;
    const int b = (tracer_id - 1) / s.kb, lane = (tracer_id - 1) % s.kb;
    const size_t bytes = (size_t)(nl - 1) * s.Np * s.kb * storage_bytes(s.prec);
    if (!s.tabbuf[b] && dalloc_bytes(s, &s.tabbuf[b], bytes)) return -1;
    if (!s.tab_on[b]) {  // the block's other lanes keep valuesAB = values
        HIPCHK(hipMemcpyAsync(s.tabbuf[b], s.dT[b], bytes, hipMemcpyDeviceToDevice, s.stream));
        s.tab_on[b] = 1;
    }
    return to_device(s, static_cast<const double *>(values), true, s.tabbuf[b], lane);
}

int tracer_get_tendencies_device(void *dttf_h, void *dttf_v, int nz, int nn) {
    State &s = S();
    if (!s.on_device || !s.acc_h) return fail("tracer_get_tendencies_device: no tendencies computed yet");
    if (nz != s.nl - 1 || nn != s.N) return fail("tracer_get_tendencies_device: expected (%d, %d)", s.nl - 1, s.N);
    if (settle_tendencies(s)) return -1;
    const double *src[2] = {s.acc_h, s.acc_v};
    void *dst[2] = {dttf_h, dttf_v};
    for (int i = 0; i < 2; ++i) {
        if (!dst[i]) continue;
        const hipError_t e = launch_relayout<double>(nullptr, nullptr, src[i], static_cast<double *>(dst[i]),
                                                     s.d_perm_n, s.N, nz, s.Np, 1, 0, 1, s.stream);
        if (e != hipSuccess) return fail("relayout launch failed: %s", hipGetErrorString(e));
    }
    HIPCHK(hipStreamSynchronize(s.stream));
    return 0;
}

int tracer_set_node_dt(const double *factor, int nod2D) {
    State &s = S();
    if (!s.dims) return fail("tracer_set_node_dt: call tracer_set_mesh_dims first");
    if (factor && nod2D != s.N) return fail("tracer_set_node_dt: nod2D=%d != %d", nod2D, s.N);
    if (factor) {
        for (int n = 0; n < nod2D; ++n)
            if (!(factor[n] > 0.0) || !std::isfinite(factor[n]))
                return fail("tracer_set_node_dt: factor of node %d is %g (must be finite and > 0)", n + 1, factor[n]);
        s.node_dt.assign(factor, factor + nod2D);
    } else {
        s.node_dt.clear();
    }
    if (s.on_device) {
        HIPCHK(hipStreamSynchronize(s.stream));
        if (upload_node_dt(s)) return -1;
        s.gen++;  // kernels and captured graphs take the pointer
    }
    return 0;
}

int tracer_set_ensemble(int nmembers) {
    State &s = S();
    if (!s.dims) return fail("tracer_set_ensemble: call tracer_set_mesh_dims first");
    if (nmembers < 1 || nmembers > kMaxKB)
        return fail("tracer_set_ensemble: nmembers=%d (1 to %d)", nmembers, kMaxKB);
    if (nmembers == s.nmem) return 0;
    for (int t = 0; t < (int)s.tr_mem.size(); ++t)
        if (s.tr_mem[t] >= nmembers)
            return fail("tracer_set_ensemble: tracer %d belongs to member %d", t + 1, s.tr_mem[t] + 1);
    if (invalidate(s)) return -1;  // the device arrays change size
    const size_t nuv = (size_t)2 * (s.nl - 1) * s.T, nw = (size_t)s.nl * s.N;
    auto sized = [](const std::vector<double> &v, size_t n) { return v.size() == n ? v : std::vector<double>(n, 0.0); };
    // new members start from the first member's velocity
    while ((int)s.muv.size() < nmembers - 1) {
        s.muv.push_back(sized(s.uv, nuv));
        s.mw.push_back(sized(s.w, nw));
        s.mwi.push_back(sized(s.wi, nw));
        s.mwe.push_back(sized(s.we, nw));
    }
    s.muv.resize(nmembers - 1);
    s.mw.resize(nmembers - 1);
    s.mwi.resize(nmembers - 1);
    s.mwe.resize(nmembers - 1);
    s.nmem = nmembers;
    s.vel_gen++;
    s.gen++;
    return 0;
}

int tracer_get_ensemble(int *nmembers) {
    if (!nmembers) return fail("tracer_get_ensemble: null pointer");
    *nmembers = S().nmem;
    return 0;
}

int tracer_set_velocity_member(int member, const double *uv, const double *w, const double *w_i,
                               const double *w_e, int nl, int nod2D, int elem2D) {
    State &s = S();
    if (member < 1 || member > s.nmem)
        return fail("tracer_set_velocity_member: member=%d, nmembers=%d (tracer_set_ensemble)", member, s.nmem);
    if (member == 1) return tracer_set_velocity_elem(uv, w, w_i, w_e, nl, nod2D, elem2D);
    if (check_dims(s, nl, nod2D, "tracer_set_velocity_member")) return -1;
    if (elem2D != s.T) return fail("tracer_set_velocity_member: elem2D=%d != %d", elem2D, s.T);
    if (!uv || !w) return fail("tracer_set_velocity_member: null pointer");
    const int k = member - 2;
    s.muv[k].assign(uv, uv + (size_t)2 * (nl - 1) * elem2D);
    s.mw[k].assign(w, w + (size_t)nl * nod2D);
    s.mwi[k] = w_i ? std::vector<double>(w_i, w_i + (size_t)nl * nod2D) : s.mw[k];
    s.mwe[k] = w_e ? std::vector<double>(w_e, w_e + (size_t)nl * nod2D) : s.mw[k];
    return refresh_velocity(s);
}

int tracer_set_tracer_member(int tracer_id, int member) {
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id %d, num_tracers=%d", tracer_id, s.ntr);
    if (member < 1 || member > s.nmem)
        return fail("tracer_set_tracer_member: member=%d, nmembers=%d (tracer_set_ensemble)", member, s.nmem);
    s.tr_mem[tracer_id - 1] = member - 1;
    return s.on_device ? upload_lane_members(s) : 0;
}

int tracer_set_kernel_timing(int enable) {
    State &s = S();
    s.kernel_timing = enable != 0;
    for (double &v : s.stage_ms) v = 0.0;
    s.stage_count = 0;
    s.qflux_ms = s.step_ms = 0.0;
    s.step_count = 0;
    return 0;
}

int tracer_get_step_times(double *qflux_ms, double *step_ms, long *count) {
    if (!qflux_ms || !step_ms || !count) return fail("tracer_get_step_times: null pointer");
    const State &s = S();
    *qflux_ms = s.step_count ? s.qflux_ms / s.step_count : 0.0;
    *step_ms = s.step_count ? s.step_ms / s.step_count : 0.0;
    *count = s.step_count;
    return 0;
}

int tracer_get_kernel_times(double *stage_ms, long *count) {
    State &s = S();
    for (int k = 0; k < 4; ++k) stage_ms[k] = s.stage_count ? s.stage_ms[k] / s.stage_count : 0.0;
    *count = s.stage_count;
    return 0;
}

int tracer_last_elapsed_ms(double *ms) {
    *ms = S().last_ms;
    return 0;
}

const char *tracer_last_error(void) { return g_err.c_str(); }

const char *tracer_version(void) { return "fesomtra 0.3 (gfx950)"; }

}  // extern "C"

#include "fesomtra_noalloc.inc"
