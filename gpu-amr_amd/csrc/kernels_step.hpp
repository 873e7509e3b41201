// kernels_step.hpp -- the leaf-patch upwind update with the halo gather fused
// in, the padded (halo-exchanged) view, and the refinement criteria.
//
// State in HBM: one buffer of n^R interior doubles per leaf slot (compact,
// 4 KiB per 8^3 patch, 2 KiB per 16^2 patch), leaves in Morton order.  Ghost
// values are not stored: each workgroup gathers the ghost layer of its patch
// from the neighbours' interiors through the device neighbour table, with the
// reference's same / finer (mean) / coarser (injection) rules, into LDS.
//
// Every per-item body is __host__ __device__; the __global__ kernels only
// distribute items over threads, so the same bodies can be run on the CPU.
#pragma once

#include "device_common.hpp"

#pragma clang fp contract(off)

namespace amrhip
{

// Per-launch step constants.  dtdx[l][d] = dt / cell_size(l)[d] and
// hs[d] = 0.5 * |v_d| are formed on the host with the same IEEE operations
// the reference performs per cell (amr_solver.hpp:333,
// AdvectionPhysics.hpp:61-65).
struct step_params
{
    double   vel[3];
    double   hs[3];
    double   dtdx[kMaxLevels][3];
    uint32_t reverse; // walk this launch's leaves backwards (sweep_item)
};

// Leaf (or run) of workgroup b in a launch over nblocks items: each XCD owns
// the same contiguous range in every launch (xcd_item), walked backwards
// when `reverse` is set.  Consecutive steps alternate the direction, so that
// a step starts on the leaves the previous one wrote last -- the lines most
// likely still in the memory-side cache (8^3 step, 32768 leaves: 57 us
// alternating against 65 us always forward) -- while a tree small enough for
// the XCDs' L2s keeps finding its leaves in the L2 that wrote them.  Items
// >= count are skipped by the caller.
__device__ inline uint32_t sweep_item(uint32_t b, uint32_t nblocks, uint32_t reverse)
{
    const uint32_t per = nblocks / kXcds, j = b / kXcds;
    return (b % kXcds) * per + (reverse ? per - 1 - j : j);
}

// Output store of a step kernel.  The non-temporal form streams the new
// state past L2 so that L2 keeps the input lines that neighbouring leaves'
// ghost gathers re-read.
template <bool NT>
__device__ inline void store_cell(double* p, double v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Step constants of a leaf: the launch's own, or, for a batch of
// concatenated trees, those of the leaf's instance.
__device__ inline const step_params& params_for(const step_params& prm, const step_params* __restrict__ table,
                                                const uint16_t* __restrict__ inst, uint32_t leaf)
{
    return table ? table[inst[leaf]] : prm;
}

// Ghost value of the cell with layout coordinates c (c[dim] lies outside
// [0, n(dim)): >= n(dim) on the positive side, < 0 on the negative side) of
// the leaf whose direction-`dir` neighbour record is (L, meta).  Restates
// halo_exchange_impl_t::{same_t, finer_t, coarser_t}
// (include/ndtree/patch_utils.hpp:315-441) in compact coordinates:
//   same    -> neighbour cell at the wrapped normal coordinate;
//   finer   -> mean of the 2^R fine cells, summed from 0.0 in row-major order;
//   coarser -> injection of the coarse cell under the contact quadrant.
// The fine cells' base coordinate along every dim d is (2 v) mod n(d), as
// the reference's device twin takes it (src/cuda/halo_exchange.cu:174-176).
// Its CPU rule reduces every dim mod n(dim) of the face normal
// (patch_utils.hpp:358), the same for square patches only; for other shapes
// that rule addresses cells outside the fine patch's interior.
template <int R>
__host__ __device__ inline double ghost_value(box<R> const& B, const double* __restrict__ u, const face_link& L,
                                              uint32_t meta, int dir, const int* c)
{
    const int  dim = dir >> 1;
    const bool pos = (dir & 1) != 0;
    const int  nrm = pos ? c[dim] - B.n(dim) : c[dim] + B.n(dim); // reference's from_idxs[dim]
    switch (rel_of(meta))
    {
        case REL_SAME:
        {
            int cc[R];
            for (int d = 0; d < R; ++d) cc[d] = (d == dim) ? nrm : c[d];
            return u[static_cast<size_t>(L.ids[0]) * B.ni() + B.lin(cc)];
        }
        case REL_COARSER:
        {
            int cc[R];
            for (int d = 0; d < R; ++d)
                cc[d] = static_cast<int>(quad_of(meta, d)) * B.half(d) + ((d == dim) ? nrm : c[d]) / 2;
            return u[static_cast<size_t>(L.ids[0]) * B.ni() + B.lin(cc)];
        }
        case REL_FINER:
        {
            int k = 0, bit = 0, base[R];
            for (int d = 0; d < R; ++d)
            {
                const int v = (d == dim) ? nrm : c[d];
                base[d]     = (v * 2) % B.n(d);
                if (d != dim)
                {
                    k |= (c[d] / B.half(d)) << bit;
                    ++bit;
                }
            }
            const double* q   = u + static_cast<size_t>(L.ids[k]) * B.ni();
            double        sum = 0.0;
            for (int o = 0; o < box<R>::F; ++o)
            {
                int cc[R];
                for (int d = 0; d < R; ++d) cc[d] = base[d] + ((o >> (R - 1 - d)) & 1);
                sum += q[B.lin(cc)];
            }
            return sum / static_cast<double>(box<R>::F);
        }
        default: return 0.0;
    }
}

template <int R, int N>
__host__ __device__ inline double ghost_value(const double* __restrict__ u, const face_link& L, uint32_t meta, int dir,
                                              const int* c)
{
    constexpr box<R> B = box<R>::cube(N);
    return ghost_value<R>(B, u, L, meta, dir, c);
}

// LDS tile of the patch plus one ghost layer: (N+2)^R doubles.
template <int R, int N>
struct tile_shape
{
    static constexpr int TP = N + 2;
    static constexpr int TS = (R == 2) ? TP * TP : TP * TP * TP;
    __host__ __device__ static int idx(const int* c)
    {
        int l = 0;
        for (int d = 0; d < R; ++d) l = l * TP + (c[d] + 1);
        return l;
    }
    // stride of layout dim j inside the tile
    __host__ __device__ static constexpr int stride(int j)
    {
        return (R == 3 && j == 0) ? TP * TP : (j == R - 1 ? 1 : TP);
    }
};

// Interior cell i of the leaf into the tile.
template <int R, int N>
__host__ __device__ inline void stage_interior(double* tile, const double* __restrict__ up, int i)
{
    int c[R];
    patch_shape<R, N>::coords(i, c);
    tile[tile_shape<R, N>::idx(c)] = up[i];
}

// Ghost item g in [0, 2R * N^(R-1)): direction g / N^(R-1), tangential
// position g % N^(R-1) in row-major order over the non-normal dims.
// Returns the direction and sets c to the (padded-frame) cell of ghost item g.
template <int R, int N>
__host__ __device__ inline int ghost_cell(int g, int* c)
{
    using S       = patch_shape<R, N>;
    const int dir = g / S::FACE, dim = dir >> 1;
    int       r   = g % S::FACE;
    for (int d = R - 1; d >= 0; --d)
    {
        if (d == dim) continue;
        c[d] = r % N;
        r /= N;
    }
    c[dim] = (dir & 1) ? N : -1;
    return dir;
}

template <int R, int N>
__host__ __device__ inline void stage_ghost(double* tile, const double* __restrict__ u, uint32_t leaf,
                                            const face_link* __restrict__ links, const uint32_t* __restrict__ metas,
                                            int g)
{
    int            c[R];
    const int      dir             = ghost_cell<R, N>(g, c);
    const size_t   e               = static_cast<size_t>(leaf) * patch_shape<R, N>::NDIR + dir;
    tile[tile_shape<R, N>::idx(c)] = ghost_value<R, N>(u, links[e], metas[e], dir, c);
}

// Index into u of the one interior cell that gives the ghost value of cell c
// when the direction-`dir` neighbour (slot nb) is same-level (copy) or
// coarser (injection): ghost_value's single-cell rules.
template <int R>
__host__ __device__ inline size_t ghost_source(box<R> const& B, int32_t nb, uint32_t meta, int dir, const int* c)
{
    const int  dim    = dir >> 1;
    const int  nrm    = (dir & 1) ? c[dim] - B.n(dim) : c[dim] + B.n(dim);
    const bool coarse = rel_of(meta) == REL_COARSER;
    int        cc[R];
    for (int d = 0; d < R; ++d)
    {
        const int v = (d == dim) ? nrm : c[d];
        cc[d]       = coarse ? static_cast<int>(quad_of(meta, d)) * B.half(d) + v / 2 : v;
    }
    return static_cast<size_t>(nb) * B.ni() + B.lin(cc);
}

template <int R, int N>
__host__ __device__ inline size_t ghost_source(int32_t nb, uint32_t meta, int dir, const int* c)
{
    constexpr box<R> B = box<R>::cube(N);
    return ghost_source<R>(B, nb, meta, dir, c);
}

// Value of cell c of `leaf` (layout coordinates, at most one of them outside
// the patch): the interior cell, or the ghost value by the face rules.
template <int R>
__host__ __device__ inline double cell_or_ghost(box<R> const& B, const double* __restrict__ u, uint32_t leaf,
                                                const face_link* __restrict__ links,
                                                const uint32_t* __restrict__ metas, const int* c)
{
    for (int d = 0; d < R; ++d)
        if (c[d] < 0 || c[d] >= B.n(d))
        {
            const int    dir = 2 * d + (c[d] >= B.n(d) ? 1 : 0);
            const size_t e   = static_cast<size_t>(leaf) * box<R>::NDIR + dir;
            return ghost_value<R>(B, u, links[e], metas[e], dir, c);
        }
    return u[static_cast<size_t>(leaf) * B.ni() + B.lin(c)];
}

// A leaf's face records: relation/quadrant word and first neighbour slot per
// direction, read with one batch of scalar loads.
template <int R>
__device__ inline void load_faces(const face_link* __restrict__ L, const uint32_t* __restrict__ M, uint32_t* m,
                                  int32_t* nb)
{
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
    {
        m[k]  = M[k];
        nb[k] = L[k].ids[0];
    }
}

// Whether none of a leaf's faces has finer neighbours (every ghost is then a
// single cell and the gathers below issue all their loads back to back).
template <int R>
__device__ inline bool single_cell_ghosts(const uint32_t* m)
{
    uint32_t finer = 0;
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) finer |= (rel_of(m[k]) == REL_FINER) ? 1u : 0u;
    return finer == 0;
}

// The six ghost values of a lane of the 3D wave layout (lane = 8 hi + lo),
// from the leaf's face records (m, nb; L, M for the finer-neighbour path):
// g[0], g[1] = -z / +z of the column (y, x) = (hi, lo); g[2], g[3] = the
// (z, x) = (hi, lo) cell of the -y / +y face; g[4], g[5] = the (z, y) =
// (hi, lo) cell of the -x / +x face (skipped unless wanted).  Unless a face
// has finer neighbours, all loads are issued back to back: a per-face switch
// on the relation makes the compiler wait for every outstanding load before
// each face.
__device__ inline void gather_ghosts_3d8(const double* __restrict__ u, const uint32_t* m, const int32_t* nb,
                                         const face_link* __restrict__ L, const uint32_t* __restrict__ M, int hi,
                                         int lo, bool want_xm, bool want_xp, double* g)
{
    const int c[6][3] = { { -1, hi, lo }, { 8, hi, lo }, { hi, -1, lo }, { hi, 8, lo }, { hi, lo, -1 }, { hi, lo, 8 } };
    if (single_cell_ghosts<3>(m))
    {
        size_t src[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) src[k] = ghost_source<3, 8>(nb[k], m[k], k, c[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) g[k] = u[src[k]];
        g[4] = want_xm ? u[src[4]] : 0.0;
        g[5] = want_xp ? u[src[5]] : 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
        g[k] = ((k == 4 && !want_xm) || (k == 5 && !want_xp)) ? 0.0 : ghost_value<3, 8>(u, L[k], M[k], k, c[k]);
}

// The ghost value a lane of the 2D wave layout gathers: cell t of face q
// (lane = 16 q + t).  Branch-free over the relation, so that the wave's
// loads are all in flight at once: every lane loads the four cells a finer
// neighbour's mean needs -- a same-level or coarser neighbour's single cell
// four times -- and keeps the mean (summed from 0.0 in row-major order, as
// ghost_value) or the first value.
__device__ inline double gather_ghost_2d16(const double* __restrict__ u, const uint32_t* ms, const int32_t* nbs,
                                           const face_link* __restrict__ L, int q, int t)
{
    using S = patch_shape<2, 16>;
    int c[2];
    c[0] = (q == 0) ? -1 : (q == 1) ? 16 : t;
    c[1] = (q == 2) ? -1 : (q == 3) ? 16 : t;
    // the lane's own face record, selected from the four uniform ones
    uint32_t m  = ms[0];
    int32_t  nb = nbs[0];
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (q == k)
        {
            m  = ms[k];
            nb = nbs[k];
        }
    size_t src[4];
    if (single_cell_ghosts<2>(ms))
        return u[ghost_source<2, 16>(nb, m, q, c)];
    const bool finer = rel_of(m) == REL_FINER;
    if (finer)
    {
        // ghost_value's finer rule: fine neighbour t / 8 of the face, cells
        // (2 t) % 16 + {0, 1} along the face, the two layers next to it
        const int dim = q >> 1, nrm = (q & 1) ? c[dim] - 16 : c[dim] + 16;
        int       base[2];
        for (int d = 0; d < 2; ++d) base[d] = (((d == dim) ? nrm : c[d]) * 2) % 16;
        // second fine slot of each face, read as uniform values and selected
        int32_t id1s[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) id1s[k] = L[k].ids[1];
        int32_t id1 = id1s[0];
#pragma unroll
        for (int k = 1; k < 4; ++k)
            if (q == k) id1 = id1s[k];
        const int32_t fine = (t / S::HALF) ? id1 : nb;
#pragma unroll
        for (int o = 0; o < 4; ++o)
        {
            const int cc[2] = { base[0] + ((o >> 1) & 1), base[1] + (o & 1) };
            src[o]          = static_cast<size_t>(fine) * S::NI + S::lin(cc);
        }
    }
    else
    {
        const size_t one = ghost_source<2, 16>(nb, m, q, c);
#pragma unroll
        for (int o = 0; o < 4; ++o) src[o] = one;
    }
    double v[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) v[o] = u[src[o]];
    double sum = 0.0;
#pragma unroll
    for (int o = 0; o < 4; ++o) sum += v[o];
    return finer ? sum / static_cast<double>(S::F) : v[0];
}

// The value of lane l - 1 (row_shr) or l + 1 (row_shl) within the lane's
// 16-lane DPP row.  In both wave layouts x runs along consecutive lanes of a
// row, so this is the x neighbour; a lane whose neighbour would come from
// outside its row (or from the other x row of a 3D row) is an x boundary lane,
// which takes its ghost value instead: such lanes read 0 here (bound_ctrl),
// so the move needs no copy of the old value.  A VALU move, where __shfl
// goes through the LDS unit.  Like every cross-lane read it must run with
// all lanes active: evaluate it as a statement of its own, never as an
// operand of ?: (which would run it only in the lanes taking that branch,
// reading the others as invalid).
__device__ inline double lane_minus1(double v)
{
    const uint64_t b  = __builtin_bit_cast(uint64_t, v);
    const int      lo = static_cast<int>(b), hi = static_cast<int>(b >> 32);
    const int      l  = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xF, 0xF, true); // row_shr:1
    const int      h  = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(h)) << 32) | static_cast<uint32_t>(l));
}

// v = *p in the lanes where `take` holds, p pointing into LDS: a volatile
// LDS read inside the branch, so the compiler keeps it a ds_read under the
// lane mask (the other lanes keep v) instead of a read by every lane and a
// 64-bit select (two VALU instructions).  Used where a boundary lane
// replaces a cross-lane value with one from LDS; the LDS writes it reads
// must be ordered before it by a barrier or leaf_sync.
__device__ inline void take_lds_if(bool take, const double* p, double& v)
{
    typedef __attribute__((address_space(3))) const volatile double lds_double;
    if (__builtin_expect(take, 1)) v = *(lds_double*)(p);
}

__device__ inline double lane_plus1(double v)
{
    const uint64_t b  = __builtin_bit_cast(uint64_t, v);
    const int      lo = static_cast<int>(b), hi = static_cast<int>(b >> 32);
    const int      l  = __builtin_amdgcn_update_dpp(0, lo, 0x101, 0xF, 0xF, true); // row_shl:1
    const int      h  = __builtin_amdgcn_update_dpp(0, hi, 0x101, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(h)) << 32) | static_cast<uint32_t>(l));
}

// Upwind update of one cell from its neighbours along each physical axis d
// (lo[d] = u_L, hi[d] = u_R): the expression order of
// amr_solver::time_step_cpu (amr_solver.hpp:302-347) and
// AdvectionPhysics::rusanovFlux (AdvectionPhysics.hpp:45-66):
//   total = 0; total -= dt/dx_d * (F(u, u_R) - F(u_L, u)) for d = x, y(, z);
//   F(a, b) = 0.5 * (a v + b v) - (0.5 |v|) (b - a); out = u + total.
__host__ __device__ inline double upwind_flux(double a, double b, double v, double hs)
{
    return 0.5 * (a * v + b * v) - hs * (b - a);
}

// The update from the cell's face fluxes, fL[d] = F(u_L, u), fR[d] = F(u, u_R).
// A face flux is the same bits whether computed for the cell on its left or
// on its right, so kernels may evaluate each face once.
template <int R>
__host__ __device__ inline double update_from_fluxes(double uc, const double* fL, const double* fR, const double* dtdx)
{
    double total = 0.0;
    for (int d = 0; d < R; ++d) total -= dtdx[d] * (fR[d] - fL[d]);
    return uc + total;
}

template <int R>
__host__ __device__ inline double upwind_update(double uc, const double* lo, const double* hi, const double* dtdx,
                                                const double* vel, const double* hs)
{
    double fL[R], fR[R];
    for (int d = 0; d < R; ++d)
    {
        fL[d] = upwind_flux(lo[d], uc, vel[d], hs[d]);
        fR[d] = upwind_flux(uc, hi[d], vel[d], hs[d]);
    }
    return update_from_fluxes<R>(uc, fL, fR, dtdx);
}

// Upwind update of interior cell i from the staged tile.
template <int R, int N>
__host__ __device__ inline double upwind_cell(const double* tile, int i, const double* dtdx, const double* vel,
                                              const double* hs)
{
    using T = tile_shape<R, N>;
    int c[R];
    patch_shape<R, N>::coords(i, c);
    const int t = T::idx(c);
    double    lo[R], hi[R];
    for (int d = 0; d < R; ++d)
    {
        const int s = T::stride(R - 1 - d); // physical axis d
        lo[d]       = tile[t - s];
        hi[d]       = tile[t + s];
    }
    return upwind_update<R>(tile[t], lo, hi, dtdx, vel, hs);
}

// max over layout dims of max(|u[c+s]-u[c]|, |u[c]-u[c-s]|) for cell i.
template <int R, int N>
__host__ __device__ inline double gradient_cell(const double* tile, int i)
{
    using T = tile_shape<R, N>;
    int c[R];
    patch_shape<R, N>::coords(i, c);
    const int t = T::idx(c);
    double    g = 0.0;
    for (int j = 0; j < R; ++j)
    {
        const int s = T::stride(j);
        g           = fmax(g, fmax(fabs(tile[t + s] - tile[t]), fabs(tile[t] - tile[t - s])));
    }
    return g;
}

// Cell i of the padded view ((n(d) + 2h) per dim, row-major): interior copy,
// face halo gathered; *kind = 0 interior, 1 face halo, 2 edge/corner (value 0).
template <int R>
__host__ __device__ inline double padded_value(box<R> const& B, int h, const double* __restrict__ u, uint32_t leaf,
                                               const face_link* __restrict__ links,
                                               const uint32_t* __restrict__ metas, int i, int* kind)
{
    int c[R], r = i, outside = 0, odim = 0;
    for (int d = R - 1; d >= 0; --d)
    {
        const int p = B.n(d) + 2 * h;
        c[d]        = r % p - h;
        r /= p;
        if (c[d] < 0 || c[d] >= B.n(d))
        {
            ++outside;
            odim = d;
        }
    }
    *kind = outside < 2 ? outside : 2;
    if (outside == 0) return u[static_cast<size_t>(leaf) * B.ni() + B.lin(c)];
    if (outside == 1)
    {
        const int    dir = 2 * odim + (c[odim] >= B.n(odim) ? 1 : 0);
        const size_t e   = static_cast<size_t>(leaf) * box<R>::NDIR + dir;
        return ghost_value<R>(B, u, links[e], metas[e], dir, c);
    }
    return 0.0;
}

template <int R, int N, int H>
__host__ __device__ inline double padded_value(const double* __restrict__ u, uint32_t leaf,
                                               const face_link* __restrict__ links,
                                               const uint32_t* __restrict__ metas, int i, int* kind)
{
    constexpr box<R> B = box<R>::cube(N);
    return padded_value<R>(B, H, u, leaf, links, metas, i, kind);
}

// The upwind update of interior cell i of `leaf` with its neighbours read
// from memory (interior cells, or ghost values at the patch boundary): the
// same operands as upwind_cell on a staged tile, for patches of any shape.
template <int R>
__host__ __device__ inline double upwind_cell_direct(box<R> const& B, const double* __restrict__ u, uint32_t leaf,
                                                     const face_link* __restrict__ links,
                                                     const uint32_t* __restrict__ metas, int i, const double* dtdx,
                                                     const double* vel, const double* hs)
{
    int c[R];
    B.coords(i, c);
    double lo[R], hi[R];
    for (int d = 0; d < R; ++d)
    {
        const int j = R - 1 - d; // layout dim of physical axis d
        c[j] -= 1;
        lo[d] = cell_or_ghost<R>(B, u, leaf, links, metas, c);
        c[j] += 2;
        hi[d] = cell_or_ghost<R>(B, u, leaf, links, metas, c);
        c[j] -= 1;
    }
    return upwind_update<R>(u[static_cast<size_t>(leaf) * B.ni() + i], lo, hi, dtdx, vel, hs);
}

// gradient_cell with the neighbours read from memory (any patch shape).
template <int R>
__host__ __device__ inline double gradient_cell_direct(box<R> const& B, const double* __restrict__ u, uint32_t leaf,
                                                       const face_link* __restrict__ links,
                                                       const uint32_t* __restrict__ metas, int i)
{
    int c[R];
    B.coords(i, c);
    const double uc = u[static_cast<size_t>(leaf) * B.ni() + i];
    double       g  = 0.0;
    for (int j = 0; j < R; ++j)
    {
        c[j] += 1;
        const double up = cell_or_ghost<R>(B, u, leaf, links, metas, c);
        c[j] -= 2;
        const double um = cell_or_ghost<R>(B, u, leaf, links, metas, c);
        c[j] += 1;
        g = fmax(g, fmax(fabs(up - uc), fabs(uc - um)));
    }
    return g;
}

__host__ __device__ inline int8_t decide(double value, int level, double rthr, double cthr, int min_level,
                                         int max_level)
{
    if (value > rthr && level < max_level) return 1;
    if (value < cthr && level > min_level) return 2;
    return 0;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <int R, int N, int NT>
__device__ inline void stage_tile(double* tile, const double* __restrict__ u, uint32_t leaf,
                                  const face_link* __restrict__ links, const uint32_t* __restrict__ metas)
{
    using S          = patch_shape<R, N>;
    const double* up = u + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT) stage_interior<R, N>(tile, up, i);
    for (int g = threadIdx.x; g < S::NDIR * S::FACE; g += NT) stage_ghost<R, N>(tile, u, leaf, links, metas, g);
}

// One time step for every leaf; one workgroup per leaf patch
// (XCD-contiguous ranges).
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_step(const double* __restrict__ u, double* __restrict__ un,
                                             const uint64_t* __restrict__ keys,
                                             const face_link* __restrict__ links,
                                             const uint32_t* __restrict__ metas, step_params prm,
                                             const step_params* __restrict__ table,
                                             const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    using S = patch_shape<R, N>;
    __shared__ double tile[tile_shape<R, N>::TS];
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    stage_tile<R, N, NT>(tile, u, leaf, links, metas);
    __syncthreads();
    const step_params& P  = params_for(prm, table, inst, leaf);
    const int          lv = key_level(keys[leaf]);
    double             dtdx[R], vel[R], hs[R];
    for (int d = 0; d < R; ++d)
    {
        dtdx[d] = P.dtdx[lv][d];
        vel[d]  = P.vel[d];
        hs[d]   = P.hs[d];
    }
    double* out = un + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT) out[i] = upwind_cell<R, N>(tile, i, dtdx, vel, hs);
}

// 3D 8^3 patches, one wave (64 lanes) per leaf.  Lane l owns the z-column at
// (y, x) = (l / 8, l % 8) in registers: x- and y-neighbours come from lanes
// l -+ 1 and l -+ 8, z-neighbours from the lane's own registers.  The six
// ghost faces are gathered one face cell per lane; the -+z faces land in the
// lane that needs them, the -+y / -+x faces go through 2 KiB of LDS.  With
// no interior tile in LDS, 32 leaves are in flight per CU.
template <bool NT>
__global__ __launch_bounds__(64) void k_step_3d8(const double* __restrict__ u, double* __restrict__ un,
                                                 const uint64_t* __restrict__ keys,
                                                 const face_link* __restrict__ links,
                                                 const uint32_t* __restrict__ metas, step_params prm,
                                                 const step_params* __restrict__ table,
                                                 const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    constexpr int NI = 512;
    __shared__ double gface[4][64]; // -y, +y (index z*8+x), -x, +x (index z*8+y)
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    const int     lane = threadIdx.x, hi3 = lane >> 3, lo3 = lane & 7;
    const double* up   = u + static_cast<size_t>(leaf) * NI;
    double        r[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) r[z] = up[z * 64 + lane];
    const face_link* L = links + static_cast<size_t>(leaf) * 6;
    const uint32_t*  M = metas + static_cast<size_t>(leaf) * 6;
    uint32_t         m[6];
    int32_t          nb[6];
    load_faces<3>(L, M, m, nb);
    double g[6];
    gather_ghosts_3d8(u, m, nb, L, M, hi3, lo3, true, true, g);
    const double gzm = g[0], gzp = g[1];
#pragma unroll
    for (int k = 0; k < 4; ++k) gface[k][lane] = g[2 + k];
    __syncthreads();
    const step_params& P       = params_for(prm, table, inst, leaf);
    const int          lv      = key_level(keys[leaf]);
    const double       dtdx[3] = { P.dtdx[lv][0], P.dtdx[lv][1], P.dtdx[lv][2] };
    const int    y = hi3, x = lo3;
    double*      out = un + static_cast<size_t>(leaf) * NI;
    // z faces are evaluated once: the +z flux of cell z is the -z flux of z+1
    double fz = upwind_flux(gzm, r[0], P.vel[2], P.hs[2]);
#pragma unroll
    for (int z = 0; z < 8; ++z)
    {
        const double uc = r[z];
        const double xl = lane_minus1(uc);
        const double xr = lane_plus1(uc);
        const double yl = __shfl_up(uc, 8, 64);
        const double yr = __shfl_down(uc, 8, 64);
        const double lx = (x == 0) ? gface[2][z * 8 + y] : xl;
        const double hx = (x == 7) ? gface[3][z * 8 + y] : xr;
        const double ly = (y == 0) ? gface[0][z * 8 + x] : yl;
        const double hy = (y == 7) ? gface[1][z * 8 + x] : yr;
        const double hz = (z == 7) ? gzp : r[z < 7 ? z + 1 : 7];
        const double fzr = upwind_flux(uc, hz, P.vel[2], P.hs[2]);
        const double fL[3] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), upwind_flux(ly, uc, P.vel[1], P.hs[1]), fz };
        const double fR[3] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), upwind_flux(uc, hy, P.vel[1], P.hs[1]), fzr };
        store_cell<NT>(out + z * 64 + lane, update_from_fluxes<3>(uc, fL, fR, dtdx));
        fz = fzr;
    }
}

// 3D 8^3 patches, one wave per run of same-level leaves along +x (leaf_run).
// The lane layout is k_step_3d8's.  Walking the run front to back, the -x
// ghosts of a leaf are the previous leaf's x = 7 column (left in LDS) and
// its +x ghosts the next leaf's x = 0 column (that leaf's interior is
// loaded before the current leaf is updated, so its loads overlap the
// update); only the run's two end faces are gathered from memory.  An x face
// touches every 64 B row of the neighbour patch, so this removes most of the
// ghost traffic that does not hit in L2.
template <bool NT>
__global__ __launch_bounds__(64) void k_step_3d8_run(const double* __restrict__ u, double* __restrict__ un,
                                                     const uint64_t* __restrict__ keys,
                                                     const face_link* __restrict__ links,
                                                     const uint32_t* __restrict__ metas, step_params prm,
                                                     const step_params* __restrict__ table,
                                                     const uint16_t* __restrict__ inst,
                                                     const leaf_run* __restrict__ runs, uint32_t nruns)
{
    constexpr int NI = 512;
    __shared__ double gface[4][64]; // -y, +y (index z*8+x), -x, +x (index z*8+y)
    const uint32_t r = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (r >= nruns) return;
    const int      lane = threadIdx.x, hi3 = lane >> 3, lo3 = lane & 7;
    const int      y = hi3, x = lo3;
    const leaf_run run  = runs[r];
    uint32_t       leaf = run.head;
    double         cur[8], nxt[8];
    {
        const double* up = u + static_cast<size_t>(leaf) * NI;
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            cur[z] = up[z * 64 + lane];
            nxt[z] = 0.0;
        }
    }
    for (uint32_t i = 0; i < run.len; ++i)
    {
        const face_link* L    = links + static_cast<size_t>(leaf) * 6;
        const uint32_t*  M    = metas + static_cast<size_t>(leaf) * 6;
        const bool       more = i + 1 < run.len;
        // ghosts of this leaf first, then the next leaf's interior, so that
        // the update below waits only for the ghosts
        uint32_t m[6];
        int32_t  nb[6];
        load_faces<3>(L, M, m, nb);
        double g[6];
        gather_ghosts_3d8(u, m, nb, L, M, hi3, lo3, i == 0, !more, g);
        const double gzm = g[0], gzp = g[1];
        gface[0][lane]   = g[2];
        gface[1][lane]   = g[3];
        if (i == 0) gface[2][lane] = g[4];
        if (!more) gface[3][lane] = g[5];
        const uint32_t next = more ? static_cast<uint32_t>(L[5].ids[0]) : leaf;
        if (more)
        {
            const double* np = u + static_cast<size_t>(next) * NI;
#pragma unroll
            for (int z = 0; z < 8; ++z) nxt[z] = np[z * 64 + lane];
        }
        __syncthreads();
        const step_params& P       = params_for(prm, table, inst, leaf);
        const int          lv      = key_level(keys[leaf]);
        const double       dtdx[3] = { P.dtdx[lv][0], P.dtdx[lv][1], P.dtdx[lv][2] };
        double*            out     = un + static_cast<size_t>(leaf) * NI;
        double             fz      = upwind_flux(gzm, cur[0], P.vel[2], P.hs[2]); // z faces once
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            const double uc = cur[z];
            const double xl = lane_minus1(uc);
            const double xr = lane_plus1(uc);
            const double yl = __shfl_up(uc, 8, 64);
            const double yr = __shfl_down(uc, 8, 64);
            const double xn = __shfl(nxt[z], lane & ~7, 64); // next leaf, x = 0, same (z, y)
            const double lx = (x == 0) ? gface[2][z * 8 + y] : xl;
            const double hx = (x == 7) ? (more ? xn : gface[3][z * 8 + y]) : xr;
            const double ly = (y == 0) ? gface[0][z * 8 + x] : yl;
            const double hy = (y == 7) ? gface[1][z * 8 + x] : yr;
            const double hz = (z == 7) ? gzp : cur[z < 7 ? z + 1 : 7];
            const double fzr = upwind_flux(uc, hz, P.vel[2], P.hs[2]);
            const double fL[3] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), upwind_flux(ly, uc, P.vel[1], P.hs[1]), fz };
            const double fR[3] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), upwind_flux(uc, hy, P.vel[1], P.hs[1]), fzr };
            store_cell<NT>(out + z * 64 + lane, update_from_fluxes<3>(uc, fL, fR, dtdx));
            fz = fzr;
        }
        __syncthreads();
        // this leaf's x = 7 column is the next leaf's -x ghost face
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            if (x == 7) gface[2][z * 8 + y] = cur[z];
            cur[z] = nxt[z];
        }
        leaf = next;
        __syncthreads();
    }
}

// 2D 16^2 patches, one wave per leaf.  Lane l owns x = l % 16 of rows
// y = 4 (l / 16) .. 4 (l / 16) + 3; x-neighbours from lanes l -+ 1, the row
// above / below a lane's four from lanes l -+ 16; the four ghost faces are
// gathered one face cell per lane into 512 B of LDS.
template <bool NT>
__global__ __launch_bounds__(64) void k_step_2d16(const double* __restrict__ u, double* __restrict__ un,
                                                  const uint64_t* __restrict__ keys,
                                                  const face_link* __restrict__ links,
                                                  const uint32_t* __restrict__ metas, step_params prm,
                                                  const step_params* __restrict__ table,
                                                  const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    constexpr int NI = 256;
    __shared__ double gface[4][16]; // -y, +y (index x), -x, +x (index y)
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    const int     lane = threadIdx.x, x = lane & 15, q = lane >> 4;
    const double* up   = u + static_cast<size_t>(leaf) * NI;
    double        r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = up[(4 * q + j) * 16 + x];
    const face_link* L = links + static_cast<size_t>(leaf) * 4;
    const uint32_t*  M = metas + static_cast<size_t>(leaf) * 4;
    uint32_t         m[4];
    int32_t          nb[4];
    load_faces<2>(L, M, m, nb);
    // lane (q, x) gathers cell x of face q
    gface[q][x] = gather_ghost_2d16(u, m, nb, L, q, x);
    __syncthreads();
    const step_params& P       = params_for(prm, table, inst, leaf);
    const int          lv      = key_level(keys[leaf]);
    const double       dtdx[2] = { P.dtdx[lv][0], P.dtdx[lv][1] };
    const double above   = __shfl_up(r[3], 16, 64);  // row 4q - 1
    const double below   = __shfl_down(r[0], 16, 64); // row 4q + 4
    double*      out     = un + static_cast<size_t>(leaf) * NI;
    // y faces are evaluated once: the +y flux of row j is the -y flux of j+1
    double fy = upwind_flux((q == 0) ? gface[0][x] : above, r[0], P.vel[1], P.hs[1]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        const int    y  = 4 * q + j;
        const double uc = r[j];
        const double xl = lane_minus1(uc);
        const double xr = lane_plus1(uc);
        const double lx = (x == 0) ? gface[2][y] : xl;
        const double hx = (x == 15) ? gface[3][y] : xr;
        const double hy = (j < 3) ? r[j < 3 ? j + 1 : 3] : ((q == 3) ? gface[1][x] : below);
        const double fyr   = upwind_flux(uc, hy, P.vel[1], P.hs[1]);
        const double fL[2] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), fy };
        const double fR[2] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), fyr };
        store_cell<NT>(out + y * 16 + x, update_from_fluxes<2>(uc, fL, fR, dtdx));
        fy = fyr;
    }
}

// 2D 16^2 patches, one wave per run of same-level leaves along +x, the lane
// layout of k_step_2d16 and the hand-over of k_step_3d8_run: the -x ghosts
// of a leaf are the previous leaf's x = 15 column (left in LDS), its +x
// ghosts the next leaf's x = 0 column (registers of the prefetched leaf).
template <bool NT>
__global__ __launch_bounds__(64) void k_step_2d16_run(const double* __restrict__ u, double* __restrict__ un,
                                                      const uint64_t* __restrict__ keys,
                                                      const face_link* __restrict__ links,
                                                      const uint32_t* __restrict__ metas, step_params prm,
                                                      const step_params* __restrict__ table,
                                                      const uint16_t* __restrict__ inst,
                                                      const leaf_run* __restrict__ runs, uint32_t nruns)
{
    constexpr int NI = 256;
    __shared__ double gface[4][16]; // -y, +y (index x), -x, +x (index y)
    const uint32_t r = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (r >= nruns) return;
    const int      lane = threadIdx.x, x = lane & 15, q = lane >> 4;
    const leaf_run run  = runs[r];
    uint32_t       leaf = run.head;
    double         cur[4], nxt[4];
    {
        const double* up = u + static_cast<size_t>(leaf) * NI;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            cur[j] = up[(4 * q + j) * 16 + x];
            nxt[j] = 0.0;
        }
    }
    for (uint32_t i = 0; i < run.len; ++i)
    {
        const bool   more = i + 1 < run.len;
        const size_t e0   = static_cast<size_t>(leaf) * 4;
        // lanes of quarter q gather face q (y faces always, the run's end x faces)
        uint32_t m[4];
        int32_t  nb[4];
        load_faces<2>(links + e0, metas + e0, m, nb);
        if (q < 2 || (q == 2 && i == 0) || (q == 3 && !more))
            gface[q][x] = gather_ghost_2d16(u, m, nb, links + e0, q, x);
        const uint32_t next = more ? static_cast<uint32_t>(links[e0 + 3].ids[0]) : leaf;
        if (more)
        {
            const double* np = u + static_cast<size_t>(next) * NI;
#pragma unroll
            for (int j = 0; j < 4; ++j) nxt[j] = np[(4 * q + j) * 16 + x];
        }
        __syncthreads();
        const step_params& P       = params_for(prm, table, inst, leaf);
        const int          lv      = key_level(keys[leaf]);
        const double       dtdx[2] = { P.dtdx[lv][0], P.dtdx[lv][1] };
        const double       above   = __shfl_up(cur[3], 16, 64);  // row 4q - 1
        const double       below   = __shfl_down(cur[0], 16, 64); // row 4q + 4
        double*            out     = un + static_cast<size_t>(leaf) * NI;
        double             fy      = upwind_flux((q == 0) ? gface[0][x] : above, cur[0], P.vel[1], P.hs[1]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int    y  = 4 * q + j;
            const double uc = cur[j];
            const double xl = lane_minus1(uc);
            const double xr = lane_plus1(uc);
            const double xn = __shfl(nxt[j], lane & ~15, 64); // next leaf, x = 0, same row
            const double lx = (x == 0) ? gface[2][y] : xl;
            const double hx = (x == 15) ? (more ? xn : gface[3][y]) : xr;
            const double hy = (j < 3) ? cur[j < 3 ? j + 1 : 3] : ((q == 3) ? gface[1][x] : below);
            const double fyr   = upwind_flux(uc, hy, P.vel[1], P.hs[1]);
            const double fL[2] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), fy };
            const double fR[2] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), fyr };
            store_cell<NT>(out + y * 16 + x, update_from_fluxes<2>(uc, fL, fR, dtdx));
            fy = fyr;
        }
        __syncthreads();
        // this leaf's x = 15 column is the next leaf's -x ghost face
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            if (x == 15) gface[2][4 * q + j] = cur[j];
            cur[j] = nxt[j];
        }
        leaf = next;
        __syncthreads();
    }
}

// Padded view with H ghost layers (the reference's patch after
// halo_exchange_update): interior copied, face halos gathered, corners 0.
template <int R, int N, int H, int NT>
__global__ __launch_bounds__(NT) void k_padded_view(const double* __restrict__ u, double* __restrict__ pad,
                                                    const face_link* __restrict__ links,
                                                    const uint32_t* __restrict__ metas, uint32_t nleaves)
{
    constexpr int  P    = N + 2 * H;
    constexpr int  FLAT = (R == 2) ? P * P : P * P * P;
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    for (int i = threadIdx.x; i < FLAT; i += NT)
    {
        int kind;
        pad[static_cast<size_t>(leaf) * FLAT + i] = padded_value<R, N, H>(u, leaf, links, metas, i, &kind);
    }
}

template <int NT>
__device__ inline double block_max(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1)
    {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// Undivided-gradient criterion: one workgroup per leaf.
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_criterion_gradient(const double* __restrict__ u,
                                                           const uint64_t* __restrict__ keys,
                                                           const face_link* __restrict__ links,
                                                           const uint32_t* __restrict__ metas, double rthr,
                                                           double cthr, int min_level, int max_level,
                                                           int8_t* __restrict__ flags, uint32_t nleaves)
{
    using S = patch_shape<R, N>;
    __shared__ double tile[tile_shape<R, N>::TS];
    __shared__ double red[NT];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    stage_tile<R, N, NT>(tile, u, leaf, links, metas);
    __syncthreads();
    double g = 0.0;
    for (int i = threadIdx.x; i < S::NI; i += NT) g = fmax(g, gradient_cell<R, N>(tile, i));
    g = block_max<NT>(g, red);
    if (threadIdx.x == 0) flags[leaf] = decide(g, key_level(keys[leaf]), rthr, cthr, min_level, max_level);
}

// Face-halo cells of a patch with h ghost layers: h * face(dim) items per
// direction, direction after direction.
template <int R>
__host__ __device__ inline int halo_items(box<R> const& B, int h)
{
    int items = 0;
    for (int dir = 0; dir < box<R>::NDIR; ++dir) items += h * B.face(dir >> 1);
    return items;
}

// Halo item g (0 <= g < halo_items): its direction (returned) and its cell c
// in compact coordinates (layer `layer` beyond the face, tangential cells in
// row-major order).
template <int R>
__host__ __device__ inline int halo_item(box<R> const& B, int h, int g, int* c)
{
    int dir = 0, r = g;
    while (r >= h * B.face(dir >> 1))
    {
        r -= h * B.face(dir >> 1);
        ++dir;
    }
    const int dim = dir >> 1, fc = B.face(dim), layer = r / fc;
    r %= fc;
    for (int d = R - 1; d >= 0; --d)
    {
        if (d == dim) continue;
        c[d] = r % B.n(d);
        r /= B.n(d);
    }
    c[dim] = (dir & 1) ? B.n(dim) + layer : -1 - layer;
    return dir;
}

// Maximum over one leaf's interior and its h-deep face halos, starting from
// 0, reduced over the workgroup (every thread must call it): the interior
// with coalesced loads, then the face halos by the ghost rules (padded_value's
// cells of kind 0 and 1; the maximum is order-free).
template <int R, int NT>
__device__ inline double leaf_max_value(box<R> const& B, int h, const double* __restrict__ u, uint32_t leaf,
                                        const face_link* __restrict__ links, const uint32_t* __restrict__ metas,
                                        double* red)
{
    double        m  = 0.0;
    const int     ni = B.ni();
    const double* up = u + static_cast<size_t>(leaf) * ni;
    for (int i = threadIdx.x; i < ni; i += NT) m = fmax(m, up[i]);
    const int items = halo_items<R>(B, h);
    for (int g = threadIdx.x; g < items; g += NT)
    {
        int            c[R];
        const int      dir = halo_item<R>(B, h, g, c);
        const size_t   e   = static_cast<size_t>(leaf) * box<R>::NDIR + dir;
        m                  = fmax(m, ghost_value<R>(B, u, links[e], metas[e], dir, c));
    }
    return block_max<NT>(m, red);
}

// Maximum-value criterion over the interior and the H-deep face halos,
// starting from 0.
template <int R, int N, int H, int NT>
__global__ __launch_bounds__(NT) void k_criterion_max(const double* __restrict__ u,
                                                      const uint64_t* __restrict__ keys,
                                                      const face_link* __restrict__ links,
                                                      const uint32_t* __restrict__ metas, double rthr, double cthr,
                                                      int min_level, int max_level, int8_t* __restrict__ flags,
                                                      uint32_t nleaves)
{
    constexpr box<R>  B = box<R>::cube(N);
    __shared__ double red[NT];
    const uint32_t    leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    const double m = leaf_max_value<R, NT>(B, H, u, leaf, links, metas, red);
    if (threadIdx.x == 0) flags[leaf] = decide(m, key_level(keys[leaf]), rthr, cthr, min_level, max_level);
}

// ---------------------------------------------------------------------------
// Kernels for patches of any shape (the box is a launch argument): one
// workgroup per leaf, neighbour cells read from memory -- the operands and
// expressions of the fixed-shape kernels, so the results are the same bits.
// ---------------------------------------------------------------------------

template <int R, int NT>
__global__ __launch_bounds__(NT) void k_step_box(box<R> B, const double* __restrict__ u, double* __restrict__ un,
                                                 const uint64_t* __restrict__ keys,
                                                 const face_link* __restrict__ links,
                                                 const uint32_t* __restrict__ metas, step_params prm,
                                                 const step_params* __restrict__ table,
                                                 const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    const step_params& P  = params_for(prm, table, inst, leaf);
    const int          lv = key_level(keys[leaf]);
    double             dtdx[R], vel[R], hs[R];
    for (int d = 0; d < R; ++d)
    {
        dtdx[d] = P.dtdx[lv][d];
        vel[d]  = P.vel[d];
        hs[d]   = P.hs[d];
    }
    const int ni  = B.ni();
    double*   out = un + static_cast<size_t>(leaf) * ni;
    for (int i = threadIdx.x; i < ni; i += NT)
        out[i] = upwind_cell_direct<R>(B, u, leaf, links, metas, i, dtdx, vel, hs);
}

template <int R, int NT>
__global__ __launch_bounds__(NT) void k_padded_box(box<R> B, int h, const double* __restrict__ u,
                                                   double* __restrict__ pad, const face_link* __restrict__ links,
                                                   const uint32_t* __restrict__ metas, uint32_t nleaves)
{
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    const int flat = B.padded(h);
    for (int i = threadIdx.x; i < flat; i += NT)
    {
        int kind;
        pad[static_cast<size_t>(leaf) * flat + i] = padded_value<R>(B, h, u, leaf, links, metas, i, &kind);
    }
}

template <int R, int NT>
__global__ __launch_bounds__(NT) void k_criterion_gradient_box(box<R> B, const double* __restrict__ u,
                                                               const uint64_t* __restrict__ keys,
                                                               const face_link* __restrict__ links,
                                                               const uint32_t* __restrict__ metas, double rthr,
                                                               double cthr, int min_level, int max_level,
                                                               int8_t* __restrict__ flags, uint32_t nleaves)
{
    __shared__ double red[NT];
    const uint32_t    leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    double    g  = 0.0;
    const int ni = B.ni();
    for (int i = threadIdx.x; i < ni; i += NT) g = fmax(g, gradient_cell_direct<R>(B, u, leaf, links, metas, i));
    g = block_max<NT>(g, red);
    if (threadIdx.x == 0) flags[leaf] = decide(g, key_level(keys[leaf]), rthr, cthr, min_level, max_level);
}

template <int R, int NT>
__global__ __launch_bounds__(NT) void k_criterion_max_box(box<R> B, int h, const double* __restrict__ u,
                                                          const uint64_t* __restrict__ keys,
                                                          const face_link* __restrict__ links,
                                                          const uint32_t* __restrict__ metas, double rthr, double cthr,
                                                          int min_level, int max_level, int8_t* __restrict__ flags,
                                                          uint32_t nleaves)
{
    __shared__ double red[NT];
    const uint32_t    leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    const double m = leaf_max_value<R, NT>(B, h, u, leaf, links, metas, red);
    if (threadIdx.x == 0) flags[leaf] = decide(m, key_level(keys[leaf]), rthr, cthr, min_level, max_level);
}

} // namespace amrhip
