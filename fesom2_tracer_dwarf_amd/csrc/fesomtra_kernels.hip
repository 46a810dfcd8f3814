// gfx950 kernels of the FP64 tracer advection step and their launchers.
//
// One tracer step with the FCT limiter is four kernels (do_oce_adv_tra,
// lib/oce_adv_tra_driver.F90:23-294):
//   k_edge    per (edge, level): low-order upwind flux and high-order minus low-order flux,
//             MUSCL/MFCT gradients reconstructed on the fly (no edge_up_dn_grad array)
//   k_node_lo per (node, level): deterministic CSR gather of the upwind fluxes, vertical upwind
//             flux, low-order solution fct_LO, vertical antidiffusive flux (QR4C/PPM/CDIFF/UPW1)
//   k_bounds  per column tile:   2-ring cluster bounds, vertical widening in LDS, P+/P- gathers,
//             limiter ratios R+/R- (lib/oce_adv_tra_fct.F90:96-374)
//   k_final   per (node, level): limiting of both flux families with R at both ends,
//             flux divergence, tendencies and the tracer update (in place)
// Without FCT: k_edge, k_vert (vertical high-order flux) and k_final.
//
// No atomics: every node sums its incident edges in the reference's edge order, so the result
// is deterministic and equal to the serial reference.
#include "fesomtra_device.hpp"
#include "fesomtra_kernels.hpp"

namespace fesomtra {

// ---------------------------------------------------------------------------------------------
template <int HOR, bool FCT>
__global__ __launch_bounds__(256) void k_edge(MeshDev m, int cols, const double *__restrict__ T,
                                              const double *__restrict__ TAB, double num_ord,
                                              double *__restrict__ flo, double *__restrict__ adfh) {
    const Tile t = tile_of(m.L, cols, m.E);
    if (!t.live) return;
    const int e = t.ent, nz = t.k + 1, L = m.L;
    const EdgeGeo g = edge_geo(m, e);
    const int nu12u = g.nu2 > 0 ? min(g.nu1, g.nu2) : g.nu1;
    const int nl12u = max(g.nl1, g.nl2);
    if (nz < nu12u || nz > nl12u) return;
    const double v = edge_vflux(m, g, e, nz);
    double fin = 0.0;
    if (FCT) {
        fin = upw1_edge(T[(size_t)g.n1 * L + t.k], T[(size_t)g.n2 * L + t.k], v);
        flo[(size_t)e * L + t.k] = fin;
    }
    const double a1 = TAB[(size_t)g.n1 * L + t.k], a2 = TAB[(size_t)g.n2 * L + t.k];
    double f;
    if (HOR == HOR_UPW1) {
        f = upw1_edge(a1, a2, v) - fin;
    } else {
        double gr[4];
        edge_grad(m, TAB, g, e, nz, gr);
        f = ho_edge<HOR>(m, g, e, nz, a1, a2, gr, v, num_ord, fin);
    }
    adfh[(size_t)e * L + t.k] = f;
}

// ---------------------------------------------------------------------------------------------
// Low-order solution and vertical antidiffusive flux (lib/oce_adv_tra_driver.F90:102-216 and the
// high-order vertical call :266-282 with o_init_zero=.false.).
template <int VER>
__global__ __launch_bounds__(256) void k_node_lo(MeshDev m, int cols, const double *__restrict__ T,
                                                 const double *__restrict__ TAB, double num_ord,
                                                 double dt, const double *__restrict__ flo,
                                                 double *__restrict__ LO, double *__restrict__ adfv) {
    const Tile t = tile_of(m.L, cols, m.N);
    if (!t.live) return;
    const int n = t.ent, nz = t.k + 1, L = m.L, NL = m.nl;
    const int nzmin = m.ulev_n[n], nzmax = m.nlev_n[n];
    if (nz < nzmin || nz > nzmax - 1) return;
    double s = 0.0;
    const int b = m.ne_ptr[n], eend = m.ne_ptr[n + 1];
    for (int i = b; i < eend; ++i) {
        const int2 q = m.ne_list[i];
        if (!in_range(q.y, nz)) continue;
        const double f = flo[(size_t)(q.x >> 1) * L + t.k];
        s = (q.x & 1) ? s - f : s + f;
    }
    const Column cl = column(m, T, m.we, n);
    const double fv0 = ver_upw1(cl, nz, 0.0);
    const double fv1 = ver_upw1(cl, nz + 1, 0.0);
    const size_t ix = (size_t)n * L + t.k;
    LO[ix] = (T[ix] * m.hnode[ix] + (s + (fv0 - fv1)) * dt / m.areasvol[(size_t)n * NL + t.k]) /
             m.hnode_new[ix];
    const Column ch = column(m, TAB, m.w, n);
    adfv[(size_t)n * NL + t.k] = ver_ho<VER>(ch, nz, num_ord, dt, fv0);
    if (nz == nzmax - 1) adfv[(size_t)n * NL + nz] = ver_ho<VER>(ch, nzmax, num_ord, dt, fv1);
}

// Vertical high-order flux without the limiter (pwvel = we, o_init_zero=.true.).
template <int VER>
__global__ __launch_bounds__(256) void k_vert(MeshDev m, int cols, const double *__restrict__ TAB,
                                              double num_ord, double dt, double *__restrict__ fv) {
    const Tile t = tile_of(m.L, cols, m.N);
    if (!t.live) return;
    const int n = t.ent, nz = t.k + 1, NL = m.nl;
    const int nzmin = m.ulev_n[n], nzmax = m.nlev_n[n];
    if (nz < nzmin || nz > nzmax - 1) return;
    const Column ch = column(m, TAB, m.we, n);
    fv[(size_t)n * NL + t.k] = ver_ho<VER>(ch, nz, num_ord, dt, 0.0);
    if (nz == nzmax - 1) fv[(size_t)n * NL + nz] = ver_ho<VER>(ch, nzmax, num_ord, dt, 0.0);
}

// ---------------------------------------------------------------------------------------------
// FCT bounds and limiter ratios (lib/oce_adv_tra_fct.F90:96-374).
__global__ __launch_bounds__(256) void k_bounds(MeshDev m, int cols, const double *__restrict__ T,
                                                const double *__restrict__ LO,
                                                const double *__restrict__ adfh,
                                                const double *__restrict__ adfv, double dt,
                                                double *__restrict__ Rp, double *__restrict__ Rm) {
    __shared__ double s_max[256], s_min[256];
    const Tile t = tile_of(m.L, cols, m.N);
    const int L = m.L, NL = m.nl;
    const int n = t.ent, nz = t.k + 1;
    int nu1 = 0, nl1 = 0;
    bool wet = false;
    if (t.live) {
        nu1 = m.ulev_n[n];
        nl1 = m.nlev_n[n];
        wet = nz >= nu1 && nz <= nl1 - 1;
    }
    if (wet) {
        // a1 + a2 + a3: max / min of max(LO, T) / min(LO, T) over the nodes of every element
        // around n; elements dry at nz contribute -1e3 / +1e3.
        double vmax = 0.0, vmin = 0.0;
        const int cnt = m.nie_num[n];
        for (int i = 0; i < cnt; ++i) {
            const int el = m.nie[(size_t)n * m.max_adj + i];
            double emax, emin;
            if (nz >= m.ulev_e[el] && nz <= m.nlev_e[el] - 1) {
                const int *en = m.enodes + (size_t)el * 3;
                const size_t i0 = (size_t)en[0] * L + t.k, i1 = (size_t)en[1] * L + t.k,
                             i2 = (size_t)en[2] * L + t.k;
                const double l0 = LO[i0], l1 = LO[i1], l2 = LO[i2];
                const double t0 = T[i0], t1 = T[i1], t2 = T[i2];
                emax = dmax(dmax(dmax(l0, t0), dmax(l1, t1)), dmax(l2, t2));
                emin = dmin(dmin(dmin(l0, t0), dmin(l1, t1)), dmin(l2, t2));
            } else {
                emax = -1.0e3;
                emin = 1.0e3;
            }
            if (i == 0) {
                vmax = emax;
                vmin = emin;
            } else {
                vmax = dmax(vmax, emax);
                vmin = dmin(vmin, emin);
            }
        }
        s_max[threadIdx.x] = vmax;
        s_min[threadIdx.x] = vmin;
    }
    __syncthreads();
    if (!wet) return;
    const size_t ix = (size_t)n * L + t.k;
    const double lo = LO[ix];
    double qmax, qmin;
    const int tid = threadIdx.x;
    if (nz == nl1 - 1 || nz == nu1) {
        qmax = s_max[tid] - lo;
        qmin = s_min[tid] - lo;
    } else {
        qmax = dmax(dmax(s_max[tid - 1], s_max[tid]), s_max[tid + 1]) - lo;
        qmin = dmin(dmin(s_min[tid - 1], s_min[tid]), s_min[tid + 1]) - lo;
    }
    // b1: vertical then horizontal contributions, in the reference's order.
    const double av0 = adfv[(size_t)n * NL + t.k], av1 = adfv[(size_t)n * NL + t.k + 1];
    double pp = 0.0 + (dmax(0.0, av0) + dmax(0.0, -av1));
    double pm = 0.0 + (dmin(0.0, av0) + dmin(0.0, -av1));
    const int b = m.ne_ptr[n], eend = m.ne_ptr[n + 1];
    for (int i = b; i < eend; ++i) {
        const int2 q = m.ne_list[i];
        if (!in_range(q.y, nz)) continue;
        double a = adfh[(size_t)(q.x >> 1) * L + t.k];
        if (q.x & 1) a = -a;
        pp = pp + dmax(0.0, a);
        pm = pm + dmin(0.0, a);
    }
    // b2
    const double V = m.areasvol[(size_t)n * NL + t.k];
    const double fp = pp * dt / V + 1e-16;
    const double fm = pm * dt / V - 1e-16;
    Rp[ix] = dmin(1.0, qmax / fp);
    Rm[ix] = dmin(1.0, qmin / fm);
}

// ---------------------------------------------------------------------------------------------
// Limited vertical antidiffusive flux at interface j of node n (lib/oce_adv_tra_fct.F90:396-424).
__device__ __forceinline__ double limit_v(const double *__restrict__ adfv,
                                          const double *__restrict__ Rp,
                                          const double *__restrict__ Rm, int L, int NL, int n,
                                          int j, int nu1, int nl1) {
    const double a = adfv[(size_t)n * NL + j - 1];
    if (j > nl1 - 1) return a;  // bottom interface: never limited (always zero)
    double ae = 1.0;
    const size_t r = (size_t)n * L + j - 1;
    if (j == nu1) {
        ae = a >= 0.0 ? dmin(ae, Rp[r]) : dmin(ae, Rm[r]);
    } else if (a >= 0.0) {
        ae = dmin(ae, Rm[r - 1]);
        ae = dmin(ae, Rp[r]);
    } else {
        ae = dmin(ae, Rp[r - 1]);
        ae = dmin(ae, Rm[r]);
    }
    return ae * a;
}

// FCT limiting (b3) + oce_tra_adv_flux2dtracer (lib/oce_adv_tra_driver.F90:298-450) + caller
// update.  With acc_h/acc_v the tendencies are accumulated there (do_oce_adv_tra contract);
// otherwise values += (dttf_h + dttf_v) / hnode_new in place.  dbg_h / dbg_v, when non-null,
// receive the limited antidiffusive fluxes.
template <bool FCT>
__global__ __launch_bounds__(256) void k_final(MeshDev m, int cols, double *__restrict__ T,
                                               const double *__restrict__ LO,
                                               const double *__restrict__ fh,
                                               const double *__restrict__ fv,
                                               const double *__restrict__ Rp,
                                               const double *__restrict__ Rm, double dt,
                                               double *__restrict__ acc_h, double *__restrict__ acc_v,
                                               double *__restrict__ dbg_h, double *__restrict__ dbg_v) {
    const Tile t = tile_of(m.L, cols, m.N);
    if (!t.live) return;
    const int n = t.ent, nz = t.k + 1, L = m.L, NL = m.nl;
    const int nu1 = m.ulev_n[n], nl1 = m.nlev_n[n];
    if (nz < nu1 || nz > nl1 - 1) return;
    const size_t ix = (size_t)n * L + t.k;
    const double V = m.areasvol[(size_t)n * NL + t.k];
    const double tval = T[ix];
    double dv = acc_v ? acc_v[ix] : 0.0;
    double top, bot;
    if (FCT) {
        dv = dv - tval * m.hnode[ix] + LO[ix] * m.hnode_new[ix];
        top = limit_v(fv, Rp, Rm, L, NL, n, nz, nu1, nl1);
        bot = limit_v(fv, Rp, Rm, L, NL, n, nz + 1, nu1, nl1);
        if (dbg_v) {
            dbg_v[(size_t)n * NL + t.k] = top;
            if (nz == nl1 - 1) dbg_v[(size_t)n * NL + nz] = bot;
        }
    } else {
        top = fv[(size_t)n * NL + t.k];
        bot = fv[(size_t)n * NL + t.k + 1];
    }
    dv = dv + (top - bot) * dt / V;
    double dh = acc_h ? acc_h[ix] : 0.0;
    const int b = m.ne_ptr[n], eend = m.ne_ptr[n + 1];
    for (int i = b; i < eend; ++i) {
        const int2 q = m.ne_list[i];
        if (!in_range(q.y, nz)) continue;
        const int e = q.x >> 1;
        const bool second = q.x & 1;
        double f = fh[(size_t)e * L + t.k];
        if (FCT) {
            // R of both end nodes (lib/oce_adv_tra_fct.F90:455-468)
            const int other = second ? m.edge_ne[e].x : m.edge_ne[e].y;
            const size_t ro = (size_t)other * L + t.k;
            const int node1 = second ? other : n;
            const int node2 = second ? n : other;
            const size_t r1 = node1 == n ? ix : ro, r2 = node2 == n ? ix : ro;
            double ae = 1.0;
            if (f >= 0.0) {
                ae = dmin(ae, Rp[r1]);
                ae = dmin(ae, Rm[r2]);
            } else {
                ae = dmin(ae, Rm[r1]);
                ae = dmin(ae, Rp[r2]);
            }
            f = ae * f;
            if (dbg_h && !second) dbg_h[(size_t)e * L + t.k] = f;
        }
        dh = second ? dh - f * dt / V : dh + f * dt / V;
    }
    if (acc_h) {
        acc_h[ix] = dh;
        acc_v[ix] = dv;
    } else {
        T[ix] = tval + (dh + dv) / m.hnode_new[ix];
    }
}

// ---------------------------------------------------------------------------------------------
// Renumbering between reference order and device order: out[i*ld + k] = in[map[i]*ld + k].
__global__ void k_gather_rows(const double *__restrict__ in, double *__restrict__ out,
                              const int *__restrict__ map, int rows, int ld) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * ld) return;
    const size_t r = i / ld, k = i - r * ld;
    out[i] = in[(size_t)map[r] * ld + k];
}

// Stats over wet layers: per-block partial min / max / sum / volume sum.
__global__ __launch_bounds__(256) void k_stats_partial(MeshDev m, const double *__restrict__ T,
                                                       int all_levels, double4 *__restrict__ part) {
    __shared__ double smin[256], smax[256], ssum[256], svol[256];
    double mn = 1.0 / 0.0, mx = -1.0 / 0.0, sm = 0.0, sv = 0.0;
    const size_t total = (size_t)m.N * m.L;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int n = (int)(i / m.L), k = (int)(i - (size_t)n * m.L);
        const int nz = k + 1;
        const bool wet = nz >= m.ulev_n[n] && nz <= m.nlev_n[n] - 1;
        if (!all_levels && !wet) continue;
        const double v = T[i];
        mn = dmin(mn, v);
        mx = dmax(mx, v);
        sm += v;
        if (wet) sv += v * m.hnode[i] * m.areasvol[(size_t)n * m.nl + k];
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    svol[threadIdx.x] = sv;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = dmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = dmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
            svol[threadIdx.x] += svol[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_double4(smin[0], smax[0], ssum[0], svol[0]);
}

// ---------------------------------------------------------------------------------------------
// Launchers
static inline int cols_for(int L, int block) { return block / L; }

#define LAUNCH_CHECK()                                        \
    do {                                                      \
        hipError_t _e = hipGetLastError();                    \
        if (_e != hipSuccess) return _e;                      \
    } while (0)

hipError_t launch_fct_step(const MeshDev &m, const StepArgs &a, hipStream_t s) {
    const int B = 256;
    const int cols = cols_for(m.L, B);
    const dim3 ge((m.E + cols - 1) / cols), gn((m.N + cols - 1) / cols);
    const bool fct = a.fct;
#define MARK(i)                                            \
    do {                                                   \
        if (a.ev) {                                        \
            hipError_t _r = hipEventRecord(a.ev[i], s);    \
            if (_r != hipSuccess) return _r;               \
        }                                                  \
    } while (0)
    MARK(0);
    switch (a.hor * 2 + (fct ? 1 : 0)) {
    case HOR_UPW1 * 2 + 0: k_edge<HOR_UPW1, false><<<ge, B, 0, s>>>(m, cols, a.T, a.TAB, a.ph, a.flo, a.adfh); break;
    case HOR_UPW1 * 2 + 1: k_edge<HOR_UPW1, true><<<ge, B, 0, s>>>(m, cols, a.T, a.TAB, a.ph, a.flo, a.adfh); break;
    case HOR_MUSCL * 2 + 0: k_edge<HOR_MUSCL, false><<<ge, B, 0, s>>>(m, cols, a.T, a.TAB, a.ph, a.flo, a.adfh); break;
    case HOR_MUSCL * 2 + 1: k_edge<HOR_MUSCL, true><<<ge, B, 0, s>>>(m, cols, a.T, a.TAB, a.ph, a.flo, a.adfh); break;
    case HOR_MFCT * 2 + 0: k_edge<HOR_MFCT, false><<<ge, B, 0, s>>>(m, cols, a.T, a.TAB, a.ph, a.flo, a.adfh); break;
    case HOR_MFCT * 2 + 1: k_edge<HOR_MFCT, true><<<ge, B, 0, s>>>(m, cols, a.T, a.TAB, a.ph, a.flo, a.adfh); break;
    default: return hipErrorInvalidValue;
    }
    LAUNCH_CHECK();
    MARK(1);
    if (fct) {
        switch (a.ver) {
        case VER_UPW1: k_node_lo<VER_UPW1><<<gn, B, 0, s>>>(m, cols, a.T, a.TAB, a.pv, a.dt, a.flo, a.LO, a.adfv); break;
        case VER_QR4C: k_node_lo<VER_QR4C><<<gn, B, 0, s>>>(m, cols, a.T, a.TAB, a.pv, a.dt, a.flo, a.LO, a.adfv); break;
        case VER_CDIFF: k_node_lo<VER_CDIFF><<<gn, B, 0, s>>>(m, cols, a.T, a.TAB, a.pv, a.dt, a.flo, a.LO, a.adfv); break;
        case VER_PPM: k_node_lo<VER_PPM><<<gn, B, 0, s>>>(m, cols, a.T, a.TAB, a.pv, a.dt, a.flo, a.LO, a.adfv); break;
        default: return hipErrorInvalidValue;
        }
        LAUNCH_CHECK();
        MARK(2);
        k_bounds<<<gn, B, 0, s>>>(m, cols, a.T, a.LO, a.adfh, a.adfv, a.dt, a.Rp, a.Rm);
        LAUNCH_CHECK();
        MARK(3);
        k_final<true><<<gn, B, 0, s>>>(m, cols, a.T, a.LO, a.adfh, a.adfv, a.Rp, a.Rm, a.dt,
                                       a.acc_h, a.acc_v, a.dbg_h, a.dbg_v);
    } else {
        switch (a.ver) {
        case VER_UPW1: k_vert<VER_UPW1><<<gn, B, 0, s>>>(m, cols, a.TAB, a.pv, a.dt, a.adfv); break;
        case VER_QR4C: k_vert<VER_QR4C><<<gn, B, 0, s>>>(m, cols, a.TAB, a.pv, a.dt, a.adfv); break;
        case VER_CDIFF: k_vert<VER_CDIFF><<<gn, B, 0, s>>>(m, cols, a.TAB, a.pv, a.dt, a.adfv); break;
        case VER_PPM: k_vert<VER_PPM><<<gn, B, 0, s>>>(m, cols, a.TAB, a.pv, a.dt, a.adfv); break;
        default: return hipErrorInvalidValue;
        }
        LAUNCH_CHECK();
        MARK(2);
        MARK(3);
        k_final<false><<<gn, B, 0, s>>>(m, cols, a.T, nullptr, a.adfh, a.adfv, nullptr, nullptr,
                                        a.dt, a.acc_h, a.acc_v, nullptr, nullptr);
    }
    LAUNCH_CHECK();
    MARK(4);
#undef MARK
    return hipSuccess;
}

hipError_t launch_gather_rows(const double *in, double *out, const int *map, int rows, int ld,
                              hipStream_t s) {
    const size_t total = (size_t)rows * ld;
    if (total == 0) return hipSuccess;
    k_gather_rows<<<dim3((unsigned)((total + 255) / 256)), 256, 0, s>>>(in, out, map, rows, ld);
    return hipGetLastError();
}

hipError_t launch_stats(const MeshDev &m, const double *T, int all_levels, double4 *part,
                        int nblocks, hipStream_t s) {
    k_stats_partial<<<nblocks, 256, 0, s>>>(m, T, all_levels, part);
    return hipGetLastError();
}

}  // namespace fesomtra
