// kernels_fused.hpp -- two time steps in one pass over the state, for trees
// whose leaves all sit on one level.
//
// On such a tree every face relation is "same", the leaves tile the periodic
// domain in Morton order, and leaf slot i holds the leaf whose Morton index
// at that level is i: a neighbour slot is a dilated-integer increment of the
// slot, with no table lookup.  A workgroup reads its patch, a two-deep ghost
// layer and the edge cells next to it (the cells two steps of the face
// stencil can reach), advances the patch and its one-deep ghost layer by the
// first step in registers, then the patch by the second step, and writes
// only u^{n+2}: half the HBM traffic of two single steps.  Every cell value
// of the intermediate state is computed by the same expression, from the
// same operands, as the single-step kernel computes it, so two fused steps
// are bit-identical to two single steps.
#pragma once

#include "kernels_step.hpp"

#pragma clang fp contract(off)

namespace amrhip
{

// Constants of a fused pair of steps: dtdx[s][d] = dt_s / cell_size(level)[d]
// for step s = 0, 1 (the second step can be clamped to the remaining time).
struct step2_params
{
    double   vel[3];
    double   hs[3];
    double   dtdx[2][3];
    uint32_t level;   // the level of every leaf
    uint32_t reverse; // sweep direction (sweep_item)
};

// Morton slot m moved by delta in {-1, 0, +1} along the axis whose bits are
// `mask` (every R-th bit below R * level): the masked part is incremented or
// decremented as a dilated integer, wrapping periodically at the level's
// extent.
template <int DELTA>
__host__ __device__ inline uint32_t dilated_step(uint32_t m, uint32_t mask)
{
    const uint32_t part = m & mask, rest = m & ~mask;
    if constexpr (DELTA > 0) return (((part | ~mask) + 1u) & mask) | rest;
    else if constexpr (DELTA < 0) return ((part - 1u) & mask) | rest;
    else return m;
}

// Morton bit masks of the three axes of a single-level 3D tree at `level`.
struct morton_masks3
{
    uint32_t x, y, z;
    __host__ __device__ explicit morton_masks3(uint32_t level)
    {
        const uint32_t bits = 3u * level;
        const uint32_t all  = bits >= 32u ? ~0u : ((1u << bits) - 1u);
        x                   = 0x49249249u & all;
        y                   = x << 1;
        z                   = x << 2;
    }
    template <int DX, int DY, int DZ>
    __host__ __device__ uint32_t at(uint32_t m) const
    {
        return dilated_step<DZ>(dilated_step<DY>(dilated_step<DX>(m, x), y), z);
    }
};

// Face flux between cells lo and hi along an axis with velocity v,
// hs = 0.5 |v|: upwind_flux(lo, hi, ...), for the shared-face update below.
__device__ inline double face_flux(double lo, double hi, double v, double hs) { return upwind_flux(lo, hi, v, hs); }

// Orders a wave's own LDS writes before its later reads.  With one wave per
// workgroup this is the workgroup barrier; with several, each wave only
// waits for itself (a wave's LDS operations complete in issue order), so
// the waves of a workgroup meet only where they exchange data.
template <int W>
__device__ inline void leaf_sync()
{
    if constexpr (W == 1)
        __syncthreads();
    else
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
}

// Two steps of a single-level tree of 8^3 patches, one wave per leaf.
//
// Layouts.  Core: lane l = 8 y + x owns the z-column (y, x) in registers, as
// in k_step_3d8.  Face: for a y face lane l = 8 z + x, for an x face
// l = 8 z + y.  Ghost-ring cells (the cells just outside the -y, +y, -x, +x
// faces, z = 0..7) are advanced by step one in face layout, 64 cells per
// pass; the -z / +z ring is the core column extended to z = -1 and 8.
//
// Fluxes.  Inside the core each x and y face flux is evaluated once: a lane
// computes the flux through its +x (+y) face and takes its -x (-y) flux from
// lane l - 1 (l - 8).  The fluxes through the core's boundary faces come
// from the face-layout passes through LDS: step one's from the ring pass
// (the ring cell's flux toward the core), step two's from a pass over the
// new ring and core boundary values.  A face flux is the same bits whichever
// cell evaluates it, so every value equals the single-step kernel's.
//
// W = 4: one workgroup of four waves per 2 x 2 x 1 block of leaves (Morton
// slots 4b .. 4b + 3: x and y siblings).  A wave skips the ring passes of
// its two faces inside the block: the step-one flux through such a face is
// evaluated from the same two cells as the ring pass would (the ring cell's
// flux toward the core), and step two takes the ring's u^{n+1} from the
// neighbouring wave's core boundary layer in LDS, the value the ring pass
// would have computed.  Every value is therefore the same bits as with W = 1.
template <bool NT, int W = 1>
__global__ __launch_bounds__(64 * W) void k_step2_3d8_uniform(const double* __restrict__ u, double* __restrict__ un,
                                                              step2_params prm, uint32_t nleaves)
{
    static_assert(W == 1 || W == 4, "one leaf or a 2 x 2 x 1 block per workgroup");
    constexpr int NI = 512;
    __shared__ double fbound_w[W][4][64]; // core boundary-face fluxes -y, +y, -x, +x (face layout)
    __shared__ double bring_w[W][4][64];  // u^{n+1} of the ghost ring (face layout)
    __shared__ double bcore_w[W][4][64];  // u^{n+1} of the core's boundary layers (face layout)
    const uint32_t item = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (item * W >= nleaves) return;
    // wave within the workgroup, as a wave-uniform (scalar) value
    const int      wv   = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const uint32_t leaf = item * W + wv;
    double(*fbound)[64] = fbound_w[wv];
    double(*bring)[64]  = bring_w[wv];
    double(*bcore)[64]  = bcore_w[wv];
    // faces inside the 2 x 2 x 1 block (W = 4): +x or -x, +y or -y
    const int           fxi  = (W == 4) ? ((wv & 1) ? 2 : 3) : -1;
    const int           fyi  = (W == 4) ? ((wv & 2) ? 0 : 1) : -1;
    const int           lane = threadIdx.x & 63, hi = lane >> 3, lo = lane & 7;
    const bool          hlo = hi < 4, llo = lo < 4;
    const morton_masks3 mk(prm.level);
    auto                P = [&](uint32_t slot) { return u + static_cast<size_t>(slot) * NI; };
    const double*       own = P(leaf);

    // ---- loads --------------------------------------------------------------
    // The 18 neighbour slots are uniform over the wave and computed first;
    // a lane picks its side with a select, so every load is unconditional
    // and none waits behind a branch.
    const uint32_t s_ym = mk.at<0, -1, 0>(leaf), s_yp = mk.at<0, 1, 0>(leaf);
    const uint32_t s_xm = mk.at<-1, 0, 0>(leaf), s_xp = mk.at<1, 0, 0>(leaf);
    const uint32_t s_zm = mk.at<0, 0, -1>(leaf), s_zp = mk.at<0, 0, 1>(leaf);
    const uint32_t s_xmym = mk.at<-1, -1, 0>(leaf), s_xpym = mk.at<1, -1, 0>(leaf);
    const uint32_t s_xmyp = mk.at<-1, 1, 0>(leaf), s_xpyp = mk.at<1, 1, 0>(leaf);
    const uint32_t s_ymzm = mk.at<0, -1, -1>(leaf), s_ymzp = mk.at<0, -1, 1>(leaf);
    const uint32_t s_ypzm = mk.at<0, 1, -1>(leaf), s_ypzp = mk.at<0, 1, 1>(leaf);
    const uint32_t s_xmzm = mk.at<-1, 0, -1>(leaf), s_xmzp = mk.at<-1, 0, 1>(leaf);
    const uint32_t s_xpzm = mk.at<1, 0, -1>(leaf), s_xpzp = mk.at<1, 0, 1>(leaf);
    auto cell = [&](uint32_t slot, int off) { return u[static_cast<size_t>(slot) * NI + off]; };
    const int ecol = llo ? 7 : 0, erow = llo ? 7 * 8 : 0, epl = hlo ? 7 * 64 : 0;

    // Ring cells, face layout (p, q) = (hi, lo): value r, second layer o,
    // adjacent core cell c, edge cells tq (q = -1 / 8 side) and tp (z = -1 / 8).
    double r[4], o[4], c[4], tq[4], tp[4];
    // W = 4: each wave has one outer y face and one outer x face (index e = 0
    // for y, 1 for x: value, second layer, core cell and the two edge cells,
    // as above) and one inner y and x face (value and core cell only)
    double re[2], oe[2], ce[2], tqe[2], tpe[2], ri[2], ci[2];
    if constexpr (W == 4)
    {
        const int  zq  = hi * 64 + lo, zy = hi * 64 + lo * 8;
        const bool yme = fyi == 1; // the outer y face is -y
        const bool xme = fxi == 3; // the outer x face is -x
        re[0]  = cell(yme ? s_ym : s_yp, zq + (yme ? 7 * 8 : 0));
        oe[0]  = cell(yme ? s_ym : s_yp, zq + (yme ? 6 * 8 : 8));
        ce[0]  = own[zq + (yme ? 0 : 7 * 8)];
        tqe[0] = cell(yme ? (llo ? s_xmym : s_xpym) : (llo ? s_xmyp : s_xpyp), hi * 64 + (yme ? 7 * 8 : 0) + ecol);
        tpe[0] = cell(yme ? (hlo ? s_ymzm : s_ymzp) : (hlo ? s_ypzm : s_ypzp), epl + (yme ? 7 * 8 : 0) + lo);
        ri[0]  = cell(yme ? s_yp : s_ym, zq + (yme ? 0 : 7 * 8));
        ci[0]  = own[zq + (yme ? 7 * 8 : 0)];
        const double2 px = *reinterpret_cast<const double2*>(u + static_cast<size_t>(xme ? s_xm : s_xp) * NI + zy +
                                                             (xme ? 6 : 0));
        re[1]  = xme ? px.y : px.x;
        oe[1]  = xme ? px.x : px.y;
        ce[1]  = own[zy + (xme ? 0 : 7)];
        tqe[1] = cell(xme ? (llo ? s_xmym : s_xmyp) : (llo ? s_xpym : s_xpyp), hi * 64 + erow + (xme ? 7 : 0));
        tpe[1] = cell(xme ? (hlo ? s_xmzm : s_xmzp) : (hlo ? s_xpzm : s_xpzp), epl + lo * 8 + (xme ? 7 : 0));
        ri[1]  = cell(xme ? s_xp : s_xm, zy + (xme ? 0 : 7));
        ci[1]  = own[zy + (xme ? 7 : 0)];
    }
    else
    {
        const int zq = hi * 64 + lo; // (z, _, x) of a y face, without the row
        r[0] = cell(s_ym, zq + 7 * 8);
        c[0] = own[zq];
        r[1] = cell(s_yp, zq);
        c[1] = own[zq + 7 * 8];
        // y faces: second layer; x-side edge in leaf (-1 or +1, -+1, 0) at
        // column 7 or 0, z-side edge in leaf (0, -+1, -1 or +1) at plane 7
        // or 0
        o[0]  = cell(s_ym, zq + 6 * 8);
        tq[0] = cell(llo ? s_xmym : s_xpym, hi * 64 + 7 * 8 + ecol);
        tp[0] = cell(hlo ? s_ymzm : s_ymzp, epl + 7 * 8 + lo);
        o[1]  = cell(s_yp, zq + 8);
        tq[1] = cell(llo ? s_xmyp : s_xpyp, hi * 64 + ecol);
        tp[1] = cell(hlo ? s_ypzm : s_ypzp, epl + lo);
        const int     zy = hi * 64 + lo * 8; // (z, y, _) of an x face
        const double2 pm = *reinterpret_cast<const double2*>(u + static_cast<size_t>(s_xm) * NI + zy + 6);
        const double2 pp = *reinterpret_cast<const double2*>(u + static_cast<size_t>(s_xp) * NI + zy);
        r[2]  = pm.y;
        o[2]  = pm.x;
        c[2]  = own[zy];
        r[3]  = pp.x;
        o[3]  = pp.y;
        c[3]  = own[zy + 7];
        // x faces: y-side edge in leaf (-+1, -1 or +1, 0) at row 7 or 0,
        // z-side edge in leaf (-+1, 0, -1 or +1) at plane 7 or 0
        tq[2] = cell(llo ? s_xmym : s_xmyp, hi * 64 + erow + 7);
        tp[2] = cell(hlo ? s_xmzm : s_xmzp, epl + lo * 8 + 7);
        tq[3] = cell(llo ? s_xpym : s_xpyp, hi * 64 + erow);
        tp[3] = cell(hlo ? s_xpzm : s_xpzp, epl + lo * 8);
    }
    // core column at z = -2..9, core layout (y, x) = (hi, lo)
    double a[12];
    a[0] = cell(s_zm, 6 * 64 + lane);
    a[1] = cell(s_zm, 7 * 64 + lane);
#pragma unroll
    for (int z = 0; z < 8; ++z) a[2 + z] = own[z * 64 + lane];
    a[10] = cell(s_zp, lane);
    a[11] = cell(s_zp, 64 + lane);
    // edge cells next to the z = -1 / 8 planes: ex at x = -1 / 8 (by the
    // lane's x side), ey at y = -1 / 8 (by its y side)
    const double exm = cell(llo ? s_xmzm : s_xpzm, 7 * 64 + hi * 8 + ecol);
    const double exp = cell(llo ? s_xmzp : s_xpzp, hi * 8 + ecol);
    const double eym = cell(hlo ? s_ymzm : s_ypzm, 7 * 64 + (hlo ? 7 * 8 : 0) + lo);
    const double eyp = cell(hlo ? s_ymzp : s_ypzp, (hlo ? 7 * 8 : 0) + lo);

    const double v[3] = { prm.vel[0], prm.vel[1], prm.vel[2] };
    const double h[3] = { prm.hs[0], prm.hs[1], prm.hs[2] };

    // ---- step one: ghost ring, face layout ----------------------------------
    if constexpr (W == 4)
    {
#pragma unroll
        for (int e = 0; e < 2; ++e)
        {
            const int nd = e == 0 ? 1 : 0; // physical axis of the face normal
            const int qd = 1 - nd;         // physical axis of q
            {
                // inner face: only the flux through the core's face
                const int  f   = e == 0 ? fyi : fxi;
                const bool neg = (f & 1) == 0;
                fbound[f][lane] = face_flux(neg ? ri[e] : ci[e], neg ? ci[e] : ri[e], v[nd], h[nd]);
            }
            const int    f   = e == 0 ? (fyi ^ 1) : (fxi ^ 1); // the outer face
            const bool   neg = (f & 1) == 0;                    // ring at -1 (else at 8)
            const double uc  = re[e];
            // cross-lane reads first, with every lane active
            const double ql = lane_minus1(uc), qh = lane_plus1(uc);
            const double pl = __shfl_up(uc, 8, 64), ph = __shfl_down(uc, 8, 64);
            double       lo3[3], hi3[3];
            lo3[qd] = (lo == 0) ? tqe[e] : ql;
            hi3[qd] = (lo == 7) ? tqe[e] : qh;
            lo3[nd] = neg ? oe[e] : ce[e];
            hi3[nd] = neg ? ce[e] : oe[e];
            lo3[2]  = (hi == 0) ? tpe[e] : pl;
            hi3[2]  = (hi == 7) ? tpe[e] : ph;
            double fL[3], fR[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                fL[d] = face_flux(lo3[d], uc, v[d], h[d]);
                fR[d] = face_flux(uc, hi3[d], v[d], h[d]);
            }
            bring[f][lane]  = update_from_fluxes<3>(uc, fL, fR, prm.dtdx[0]);
            fbound[f][lane] = neg ? fR[nd] : fL[nd]; // through the core's boundary face
        }
    }
    else
    {
#pragma unroll
        for (int f = 0; f < 4; ++f)
        {
            const bool   neg = (f & 1) == 0;    // ring at -1 (else at 8)
            const int    nd  = (f < 2) ? 1 : 0; // physical axis of the face normal
            const int    qd  = 1 - nd;          // physical axis of q
            const double uc  = r[f];
            // cross-lane reads first, with every lane active
            const double ql = lane_minus1(uc), qh = lane_plus1(uc);
            const double pl = __shfl_up(uc, 8, 64), ph = __shfl_down(uc, 8, 64);
            double       lo3[3], hi3[3];
            lo3[qd] = (lo == 0) ? tq[f] : ql;
            hi3[qd] = (lo == 7) ? tq[f] : qh;
            lo3[nd] = neg ? o[f] : c[f];
            hi3[nd] = neg ? c[f] : o[f];
            lo3[2]  = (hi == 0) ? tp[f] : pl;
            hi3[2]  = (hi == 7) ? tp[f] : ph;
            double fL[3], fR[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                fL[d] = face_flux(lo3[d], uc, v[d], h[d]);
                fR[d] = face_flux(uc, hi3[d], v[d], h[d]);
            }
            bring[f][lane]  = update_from_fluxes<3>(uc, fL, fR, prm.dtdx[0]);
            fbound[f][lane] = neg ? fR[nd] : fL[nd]; // through the core's boundary face
        }
    }
    leaf_sync<W>(); // this wave's ring values and fluxes

    // ---- step one: core column at z = -1..8 ---------------------------------
    const int y = hi, x = lo;
    double    b[10];
    {
        double fz = face_flux(a[0], a[1], v[2], h[2]);
#pragma unroll
        for (int k = 0; k < 10; ++k)
        {
            const double uc  = a[k + 1];
            double       fRx = face_flux(uc, lane_plus1(uc), v[0], h[0]);
            double       fLx = lane_minus1(fRx);
            double       fRy = face_flux(uc, __shfl_down(uc, 8, 64), v[1], h[1]);
            double       fLy = __shfl_up(fRy, 8, 64);
            if (k >= 1 && k <= 8)
            {
                const int     z  = k - 1;
                const double* gx = &fbound[x < 4 ? 2 : 3][z * 8 + y];
                const double* gy = &fbound[y < 4 ? 0 : 1][z * 8 + x];
                take_lds_if(x == 0, gx, fLx);
                take_lds_if(x == 7, gx, fRx);
                take_lds_if(y == 0, gy, fLy);
                take_lds_if(y == 7, gy, fRy);
            }
            else
            {
                // z = -1 / 8 plane: boundary fluxes against the edge cells
                const double ex = (k == 0) ? exm : exp, ey = (k == 0) ? eym : eyp;
                const double bx = face_flux(x < 4 ? ex : uc, x < 4 ? uc : ex, v[0], h[0]);
                const double by = face_flux(y < 4 ? ey : uc, y < 4 ? uc : ey, v[1], h[1]);
                fLx             = (x == 0) ? bx : fLx;
                fRx             = (x == 7) ? bx : fRx;
                fLy             = (y == 0) ? by : fLy;
                fRy             = (y == 7) ? by : fRy;
            }
            const double fzr   = face_flux(uc, a[k + 2], v[2], h[2]);
            const double fL[3] = { fLx, fLy, fz };
            const double fR[3] = { fRx, fRy, fzr };
            b[k]               = update_from_fluxes<3>(uc, fL, fR, prm.dtdx[0]);
            fz                 = fzr;
        }
    }
    // the core's boundary layers of u^{n+1}, to face layout
    if (x == 0 || x == 7)
    {
#pragma unroll
        for (int z = 0; z < 8; ++z) bcore[x == 0 ? 2 : 3][z * 8 + y] = b[1 + z];
    }
    if (y == 0 || y == 7)
    {
#pragma unroll
        for (int z = 0; z < 8; ++z) bcore[y == 0 ? 0 : 1][z * 8 + x] = b[1 + z];
    }
    __syncthreads(); // every wave's core boundary layers

    // ---- step two: boundary-face fluxes, face layout ------------------------
#pragma unroll
    for (int f = 0; f < 4; ++f)
    {
        const bool   neg = (f & 1) == 0;
        const int    nd  = (f < 2) ? 1 : 0;
        // inside the block, the ring's u^{n+1} is the neighbouring wave's
        // core boundary layer on the opposite face
        const double rv = (f == fxi)   ? bcore_w[wv ^ 1][f ^ 1][lane]
                          : (f == fyi) ? bcore_w[wv ^ 2][f ^ 1][lane]
                                       : bring[f][lane];
        const double cv = bcore[f][lane];
        fbound[f][lane] = neg ? face_flux(rv, cv, v[nd], h[nd]) : face_flux(cv, rv, v[nd], h[nd]);
    }
    leaf_sync<W>(); // this wave's step-two boundary fluxes

    // ---- step two: core column at z = 0..7 ----------------------------------
    double* out = un + static_cast<size_t>(leaf) * NI;
    double  fz  = face_flux(b[0], b[1], v[2], h[2]);
#pragma unroll
    for (int z = 0; z < 8; ++z)
    {
        const double uc    = b[1 + z];
        double       fRx   = face_flux(uc, lane_plus1(uc), v[0], h[0]);
        double       fLx   = lane_minus1(fRx);
        double       fRy   = face_flux(uc, __shfl_down(uc, 8, 64), v[1], h[1]);
        double       fLy   = __shfl_up(fRy, 8, 64);
        const double* gx   = &fbound[x < 4 ? 2 : 3][z * 8 + y];
        const double* gy   = &fbound[y < 4 ? 0 : 1][z * 8 + x];
        take_lds_if(x == 0, gx, fLx);
        take_lds_if(x == 7, gx, fRx);
        take_lds_if(y == 0, gy, fLy);
        take_lds_if(y == 7, gy, fRy);
        const double fzr   = face_flux(uc, b[2 + z], v[2], h[2]);
        const double fL[3] = { fLx, fLy, fz };
        const double fR[3] = { fRx, fRy, fzr };
        store_cell<NT>(out + z * 64 + lane, update_from_fluxes<3>(uc, fL, fR, prm.dtdx[1]));
        fz = fzr;
    }
}

} // namespace amrhip
