// Tiles of the fused single-pass step: host-built index tables (device numbering) and launcher.
#pragma once

#include "fesomtra_types.hpp"

namespace fesomtra {

constexpr int kTileThreads = 768;   // region nodes per tile (one thread each; 12 waves, 168 VGPRs)
constexpr int kTileElems = 1536;    // elements of a tile's region (gradients held in LDS)
constexpr int kTileDeg = 8;         // edges per node of rings 0-1 (antidiffusive fluxes in registers)
constexpr int kTileOwned = 384;     // target owned nodes per tile (shrunk where the region is too big)
constexpr int kTileEntries = 4608;  // node->edge entries of rings 0-2 (staged in LDS)
constexpr int kTileSel = 4608;      // element-list entries of rings 0-2 (staged in LDS)

// Region "slots" of a tile: owned nodes first (ring 0), then rings 1, 2, 3.  Local indices of
// slots, entries and elements are relative to the tile.  The per-tile tables are packed so that a
// tile's whole table fits in LDS next to the level buffers:
//   entry  int2 {(edge * 2 + second) << 7 | top,  oth | up << 10 | down << 21}
//          (second: the slot's node is edges(2); top: deepest level of the edge's scatter range;
//           oth: slot of the other end; up / down: tile elements of the up / down triangles, 2047
//           when absent or not needed)
//   element corners  l0 | l1 << 10 | l2 << 20 (slots of its three nodes)
struct TileMesh {
    int ntiles;            // padded to a multiple of the XCD count (empty tiles have no slots)
    const int *t_slot0, *t_nslot, *t_sp0, *t_ent0, *t_nent, *t_sel0, *t_nsel, *t_el0, *t_nel, *t_lmax;
    const int *s_node;     // device node of a slot
    const int *s_ring;     // 0..3
    const int *s_eptr;     // per slot: first entry (tile-relative); nslot + 1 values from t_sp0
    const int *s_elptr;    // per slot: first element-list entry (tile-relative); nslot + 1 from t_sp0
    const int2 *ent;       // packed entries (rings 0-2), reference edge order
    const int *sel;        // element lists (rings 0-2), nod_in_elem2D order, tile element indices
    const int *el_id;      // device element of a tile element
    const int *el_loc;     // packed corners
};

__device__ __forceinline__ int ent_e2(int2 v) { return v.x >> 7; }
__device__ __forceinline__ int ent_top(int2 v) { return v.x & 127; }
__device__ __forceinline__ int ent_oth(int2 v) { return v.y & 1023; }
__device__ __forceinline__ int ent_up(int2 v) { return (v.y >> 10) & 2047; }
__device__ __forceinline__ int ent_dn(int2 v) { return (v.y >> 21) & 2047; }

// Area-weighted mean of the LDS element gradients around a slot's node at level nz
// (lib/oce_muscl_adv.F90:394-409), in nod_in_elem2D order; b, e delimit the slot's element list
// in the LDS copy sEl, sElId the device element of each tile element.
__device__ __forceinline__ void tile_mean_grad(const MeshCol<double> &m, const int *sEl, const int *sElId, int b,
                                               int e, int nz, const double *sGx, const double *sGy, double &gx,
                                               double &gy) {
    double tvol = 0, tx = 0, ty = 0;
    for (int i = b; i < e; ++i) {
        const int le = sEl[i];
        const int el = sElId[le];
        if (cld(m.nlev_e, el) - 1 < nz) continue;
        const double ar = cld(m.elem_area, el);
        tvol = tvol + ar;
        tx = tx + sGx[le] * ar;
        ty = ty + sGy[le] * ar;
    }
    gx = tx / tvol;
    gy = ty / tvol;
}

// One FCT step of one FP64 tracer (Tin -> Tout) with horizontal / vertical schemes hor, ver.
hipError_t launch_tile_step(const MeshCol<double> &m, const TileMesh &tm, const double *Tin, double *Tout, int hor,
                            int ver, double ph, double pv, double dt, hipStream_t s);

}  // namespace fesomtra
