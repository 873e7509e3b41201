// kernel_emulation.hip -- runs the HIP kernels' per-item bodies on the CPU
// (host-only build), one "workgroup" after another, so that the device logic
// can be compared with the oracle without a GPU.  Test infrastructure only.
#include "device_common.hpp"
#include "kernels_step.hpp"
#include "kernels_tree.hpp"
#include "kernels_adjoint.hpp"
#include "kernels_fused.hpp"

#include <cstring>
#include <vector>

using namespace amrhip;

namespace
{

template <int R>
int build_table(int depth, const uint64_t* keys, uint32_t n, face_link* links, uint32_t* metas)
{
    int err = 0;
    for (uint32_t idx = 0; idx < n * 2 * R; ++idx)
        if (!table_entry<R>(keys, n, depth, idx / (2 * R), static_cast<int>(idx % (2 * R)), &links[idx], &metas[idx]))
            err = 1;
    return err;
}

template <int R, int N>
void step(const uint64_t* keys, uint32_t n, const face_link* links, const uint32_t* metas, const double* u, double* un,
          const double* dtdx_tab, const double* vel, const double* hs)
{
    using S = patch_shape<R, N>;
    std::vector<double> tile(tile_shape<R, N>::TS, 0.0);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
    {
        for (int i = 0; i < S::NI; ++i) stage_interior<R, N>(tile.data(), u + static_cast<size_t>(leaf) * S::NI, i);
        for (int g = 0; g < S::NDIR * S::FACE; ++g) stage_ghost<R, N>(tile.data(), u, leaf, links, metas, g);
        const int lv = key_level(keys[leaf]);
        double    dtdx[3];
        for (int d = 0; d < R; ++d) dtdx[d] = dtdx_tab[lv * 3 + d];
        for (int i = 0; i < S::NI; ++i)
            un[static_cast<size_t>(leaf) * S::NI + i] = upwind_cell<R, N>(tile.data(), i, dtdx, vel, hs);
    }
}

template <int R, int N, int H>
void padded(uint32_t n, const face_link* links, const uint32_t* metas, const double* u, double* pad)
{
    constexpr int P    = N + 2 * H;
    constexpr int FLAT = (R == 2) ? P * P : P * P * P;
    for (uint32_t leaf = 0; leaf < n; ++leaf)
        for (int i = 0; i < FLAT; ++i)
        {
            int kind;
            pad[static_cast<size_t>(leaf) * FLAT + i] = padded_value<R, N, H>(u, leaf, links, metas, i, &kind);
        }
}

template <int R, int N, int H>
void flags(int kind, const uint64_t* keys, uint32_t n, const face_link* links, const uint32_t* metas, const double* u,
           double rthr, double cthr, int mn, int mx, int8_t* out)
{
    using S            = patch_shape<R, N>;
    constexpr int P    = N + 2 * H;
    constexpr int FLAT = (R == 2) ? P * P : P * P * P;
    std::vector<double> tile(tile_shape<R, N>::TS, 0.0);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
    {
        double v = 0.0;
        if (kind == 1)
        {
            for (int i = 0; i < S::NI; ++i) stage_interior<R, N>(tile.data(), u + static_cast<size_t>(leaf) * S::NI, i);
            for (int g = 0; g < S::NDIR * S::FACE; ++g) stage_ghost<R, N>(tile.data(), u, leaf, links, metas, g);
            for (int i = 0; i < S::NI; ++i) v = fmax(v, gradient_cell<R, N>(tile.data(), i));
        }
        else
        {
            for (int i = 0; i < FLAT; ++i)
            {
                int          k;
                const double x = padded_value<R, N, H>(u, leaf, links, metas, i, &k);
                if (k < 2) v = fmax(v, x);
            }
        }
        out[leaf] = decide(v, key_level(keys[leaf]), rthr, cthr, mn, mx);
    }
}

// The regrid kernel sequence of amr_ctx::regrid, serialised.
template <int R, int N>
int regrid(int depth, const uint64_t* keys, uint32_t n, const int8_t* fl, const face_link* links,
           const uint32_t* metas, const double* u, int max_level, uint64_t* keys_out, double* u_out, uint32_t cap,
           uint32_t* summary, uint32_t* maps = nullptr)
{
    using S = patch_shape<R, N>;
    std::vector<uint8_t>  mark(n), head(n);
    std::vector<uint32_t> counts(n), pos(n);
    for (uint32_t i = 0; i < n; ++i) mark[i] = refine_mark(fl[i], keys[i], depth);
    for (int l = std::min(max_level, depth); l >= 1; --l)
        for (uint32_t i = 0; i < n; ++i) close_item<R>(keys, l, links, metas, mark.data(), i);
    uint32_t nc = 0, nr = 0, err = 0;
    for (uint32_t i = 0; i < n; ++i)
    {
        bool root = false;
        head[i]   = coarsen_family_head<R>(fl, keys, n, depth, links, metas, mark.data(), i, &root) ? 1 : 0;
        if (root) err |= 2u;
        nc += head[i];
    }
    uint32_t total = 0;
    for (uint32_t i = 0; i < n; ++i)
    {
        counts[i] = emit_count<R>(keys, depth, mark.data(), head.data(), i);
        nr += mark[i];
        pos[i] = total;
        total += counts[i];
    }
    summary[0] = total;
    summary[1] = nr;
    summary[2] = nc;
    summary[3] = err;
    if (maps) // rows of n: mark, head, counts, pos
        for (uint32_t i = 0; i < n; ++i)
        {
            maps[i]         = mark[i];
            maps[n + i]     = head[i];
            maps[2 * n + i] = counts[i];
            maps[3 * n + i] = pos[i];
        }
    if (err || (nr == 0 && nc == 0) || total > cap) return 0;
    for (uint32_t i = 0; i < n; ++i)
    {
        for (uint32_t j = 0; j < counts[i]; ++j)
        {
            double* out = u_out + static_cast<size_t>(pos[i] + j) * S::NI;
            for (int t = 0; t < S::NI; ++t)
            {
                if (mark[i])
                    out[t] = refine_value<R, N>(u + static_cast<size_t>(i) * S::NI, static_cast<int>(j), t);
                else if (head[i])
                    out[t] = restrict_value<R, N>(u, i, t);
                else
                    out[t] = u[static_cast<size_t>(i) * S::NI + t];
            }
            keys_out[pos[i] + j] = emit_key<R>(keys[i], depth, mark[i] != 0, head[i] != 0, static_cast<int>(j));
        }
    }
    return 1;
}

// k_adj_step, serialised: lam = transpose(step)(lam_next), vpart per leaf.
template <int R, int N>
void adj_step(const uint64_t* keys, uint32_t n, const face_link* links, const uint32_t* metas, const double* lam_next,
              const double* u, const adjoint_params& prm, double* lam, double* vpart)
{
    using S = patch_shape<R, N>;
    std::vector<double> tile(tile_shape<R, N>::TS, 0.0), h(S::NDIR * S::FACE);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
    {
        for (int i = 0; i < S::NI; ++i) stage_interior<R, N>(tile.data(), u + static_cast<size_t>(leaf) * S::NI, i);
        for (int g = 0; g < S::NDIR * S::FACE; ++g) stage_ghost<R, N>(tile.data(), u, leaf, links, metas, g);
        const face_link* L  = links + static_cast<size_t>(leaf) * S::NDIR;
        const uint32_t*  M  = metas + static_cast<size_t>(leaf) * S::NDIR;
        const int        lv = key_level(keys[leaf]);
        for (int g = 0; g < S::NDIR * S::FACE; ++g)
        {
            const int dir = g / S::FACE;
            int       c[R];
            face_coords<R, N>(g % S::FACE, dir, c);
            h[g] = adj_face_term<R, N>(lam_next, L[dir], M[dir], dir, c, lv, prm);
        }
        double dtdx[R], vp[R], vm[R], sgn[R], acc[R];
        for (int d = 0; d < R; ++d)
        {
            dtdx[d] = prm.dtdx[lv][d];
            vp[d]   = prm.vp[d];
            vm[d]   = prm.vm[d];
            sgn[d]  = prm.sgn[d];
            acc[d]  = 0.0;
        }
        const double* lt = lam_next + static_cast<size_t>(leaf) * S::NI;
        for (int i = 0; i < S::NI; ++i)
        {
            lam[static_cast<size_t>(leaf) * S::NI + i] = adj_cell<R, N>(lt, h.data(), M, i, dtdx, vp, vm);
            adj_velocity_cell<R, N>(tile.data(), i, lt[i], dtdx, sgn, acc);
        }
        for (int d = 0; d < 3; ++d) vpart[static_cast<size_t>(leaf) * 3 + d] = d < R ? acc[d] : 0.0;
    }
}

// k_adj_regrid, serialised.
template <int R, int N>
void adj_regrid(const double* lam_new, const uint8_t* mark, const uint8_t* head, const uint32_t* counts,
                const uint32_t* pos, double* lam_old, uint32_t n_old)
{
    using S = patch_shape<R, N>;
    for (uint32_t i = 0; i < n_old; ++i)
    {
        if (counts[i] == 0) continue;
        for (int t = 0; t < S::NI; ++t)
        {
            if (mark[i])
                lam_old[static_cast<size_t>(i) * S::NI + t] = adj_refine_value<R, N>(lam_new, pos[i], t);
            else if (head[i])
                for (int kid = 0; kid < S::F; ++kid)
                    lam_old[static_cast<size_t>(i + kid) * S::NI + t] =
                        adj_restrict_value<R, N>(lam_new + static_cast<size_t>(pos[i]) * S::NI, kid, t);
            else
                lam_old[static_cast<size_t>(i) * S::NI + t] = lam_new[static_cast<size_t>(pos[i]) * S::NI + t];
        }
    }
}

// ---- the any-shape (box) kernels' bodies, serialised ----------------------

template <int R>
box<R> box_of(const int* ext)
{
    return box<R>{ { ext[0], ext[1], R == 3 ? ext[2] : 1 } };
}

template <int R>
void step_box(const int* ext, const uint64_t* keys, uint32_t n, const face_link* links, const uint32_t* metas,
              const double* u, double* un, const double* dtdx_tab, const double* vel, const double* hs)
{
    const box<R> B = box_of<R>(ext);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
    {
        const int lv = key_level(keys[leaf]);
        double    dtdx[3];
        for (int d = 0; d < R; ++d) dtdx[d] = dtdx_tab[lv * 3 + d];
        for (int i = 0; i < B.ni(); ++i)
            un[static_cast<size_t>(leaf) * B.ni() + i] = upwind_cell_direct<R>(B, u, leaf, links, metas, i, dtdx, vel, hs);
    }
}

template <int R>
void padded_box(const int* ext, int h, uint32_t n, const face_link* links, const uint32_t* metas, const double* u,
                double* pad)
{
    const box<R> B    = box_of<R>(ext);
    const int    flat = B.padded(h);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
        for (int i = 0; i < flat; ++i)
        {
            int kind;
            pad[static_cast<size_t>(leaf) * flat + i] = padded_value<R>(B, h, u, leaf, links, metas, i, &kind);
        }
}

// k_criterion_gradient_box / k_criterion_max_box (leaf_max_value's interior
// and halo items).
template <int R>
void flags_box(const int* ext, int h, int kind, const uint64_t* keys, uint32_t n, const face_link* links,
               const uint32_t* metas, const double* u, double rthr, double cthr, int mn, int mx, int8_t* out)
{
    const box<R> B = box_of<R>(ext);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
    {
        double v = 0.0;
        if (kind == 1)
            for (int i = 0; i < B.ni(); ++i) v = fmax(v, gradient_cell_direct<R>(B, u, leaf, links, metas, i));
        else
        {
            for (int i = 0; i < B.ni(); ++i) v = fmax(v, u[static_cast<size_t>(leaf) * B.ni() + i]);
            for (int g = 0; g < halo_items<R>(B, h); ++g)
            {
                int          c[R];
                const int    dir = halo_item<R>(B, h, g, c);
                const size_t e   = static_cast<size_t>(leaf) * box<R>::NDIR + dir;
                v                = fmax(v, ghost_value<R>(B, u, links[e], metas[e], dir, c));
            }
        }
        out[leaf] = decide(v, key_level(keys[leaf]), rthr, cthr, mn, mx);
    }
}

// The transfer of emit_leaf for one regrid plan (maps rows: mark, head,
// counts, pos).
template <int R>
void transfer_box(const int* ext, int depth, const uint64_t* keys, uint32_t n, const uint32_t* maps, const double* u,
                  uint64_t* keys_out, double* u_out)
{
    const box<R>    B      = box_of<R>(ext);
    const int       ni     = B.ni();
    const uint32_t* mark   = maps;
    const uint32_t* head   = maps + n;
    const uint32_t* counts = maps + 2 * n;
    const uint32_t* pos    = maps + 3 * n;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < counts[i]; ++j)
        {
            double* out = u_out + static_cast<size_t>(pos[i] + j) * ni;
            for (int t = 0; t < ni; ++t)
            {
                if (mark[i])
                    out[t] = refine_value<R>(B, u + static_cast<size_t>(i) * ni, static_cast<int>(j), t);
                else if (head[i])
                    out[t] = restrict_value<R>(B, u, i, t);
                else
                    out[t] = u[static_cast<size_t>(i) * ni + t];
            }
            keys_out[pos[i] + j] = emit_key<R>(keys[i], depth, mark[i] != 0, head[i] != 0, static_cast<int>(j));
        }
}

template <int R>
void adj_step_box(const int* ext, const uint64_t* keys, uint32_t n, const face_link* links, const uint32_t* metas,
                  const double* lam_next, const double* u, const adjoint_params& prm, double* lam, double* vpart)
{
    const box<R> B = box_of<R>(ext);
    for (uint32_t leaf = 0; leaf < n; ++leaf)
    {
        const face_link* L  = links + static_cast<size_t>(leaf) * box<R>::NDIR;
        const uint32_t*  M  = metas + static_cast<size_t>(leaf) * box<R>::NDIR;
        const int        lv = key_level(keys[leaf]);
        double           dtdx[R], vp[R], vm[R], sgn[R], acc[R];
        for (int d = 0; d < R; ++d)
        {
            dtdx[d] = prm.dtdx[lv][d];
            vp[d]   = prm.vp[d];
            vm[d]   = prm.vm[d];
            sgn[d]  = prm.sgn[d];
            acc[d]  = 0.0;
        }
        for (int i = 0; i < B.ni(); ++i)
        {
            lam[static_cast<size_t>(leaf) * B.ni() + i] = adj_cell_direct<R>(B, lam_next, leaf, L, M, i, lv, prm, dtdx, vp, vm);
            adj_velocity_cell_direct<R>(B, u, leaf, links, metas, i, lam_next[static_cast<size_t>(leaf) * B.ni() + i],
                                        dtdx, sgn, acc);
        }
        for (int d = 0; d < 3; ++d) vpart[static_cast<size_t>(leaf) * 3 + d] = d < R ? acc[d] : 0.0;
    }
}

template <int R>
void adj_regrid_box(const int* ext, const double* lam_new, const uint32_t* maps, uint32_t n_old, double* lam_old)
{
    const box<R>    B      = box_of<R>(ext);
    const int       ni     = B.ni();
    const uint32_t* mark   = maps;
    const uint32_t* head   = maps + n_old;
    const uint32_t* counts = maps + 2 * n_old;
    const uint32_t* pos    = maps + 3 * n_old;
    for (uint32_t i = 0; i < n_old; ++i)
    {
        if (counts[i] == 0) continue;
        for (int t = 0; t < ni; ++t)
        {
            if (mark[i])
                lam_old[static_cast<size_t>(i) * ni + t] = adj_refine_value<R>(B, lam_new, pos[i], t);
            else if (head[i])
                for (int kid = 0; kid < box<R>::F; ++kid)
                    lam_old[static_cast<size_t>(i + kid) * ni + t] =
                        adj_restrict_value<R>(B, lam_new + static_cast<size_t>(pos[i]) * ni, kid, t);
            else
                lam_old[static_cast<size_t>(i) * ni + t] = lam_new[static_cast<size_t>(pos[i]) * ni + t];
        }
    }
}

adjoint_params make_adjoint_params(const double* dtdx, const double* vel)
{
    adjoint_params prm{};
    std::memcpy(prm.dtdx, dtdx, sizeof(prm.dtdx));
    for (int d = 0; d < 3; ++d)
    {
        prm.vp[d]  = 0.5 * (vel[d] + std::abs(vel[d]));
        prm.vm[d]  = 0.5 * (vel[d] - std::abs(vel[d]));
        prm.sgn[d] = vel[d] > 0.0 ? 1.0 : (vel[d] < 0.0 ? -1.0 : 0.0);
    }
    return prm;
}

} // namespace

#define EMU_SHAPES(X) X(2, 4, 1) X(2, 4, 2) X(2, 8, 1) X(2, 16, 1) X(2, 10, 2) X(2, 50, 2) X(3, 4, 1) X(3, 8, 1)

extern "C"
{

int emu_build_table(int rank, int depth, const uint64_t* keys, uint32_t n, void* links, uint32_t* metas)
{
    auto* L = static_cast<face_link*>(links);
    return rank == 2 ? build_table<2>(depth, keys, n, L, metas) : build_table<3>(depth, keys, n, L, metas);
}

int emu_step(int rank, int nn, const uint64_t* keys, uint32_t n, const void* links, const uint32_t* metas,
             const double* u, double* un, const double* dtdx, const double* vel, const double* hs)
{
    auto const* L = static_cast<const face_link*>(links);
#define X(R_, N_, H_)                                                   \
    if (rank == R_ && nn == N_)                                         \
    {                                                                   \
        step<R_, N_>(keys, n, L, metas, u, un, dtdx, vel, hs);          \
        return 0;                                                       \
    }
    EMU_SHAPES(X)
#undef X
    return -1;
}

int emu_padded(int rank, int nn, int h, uint32_t n, const void* links, const uint32_t* metas, const double* u,
               double* pad)
{
    auto const* L = static_cast<const face_link*>(links);
#define X(R_, N_, H_)                                  \
    if (rank == R_ && nn == N_ && h == H_)             \
    {                                                  \
        padded<R_, N_, H_>(n, L, metas, u, pad);       \
        return 0;                                      \
    }
    EMU_SHAPES(X)
#undef X
    return -1;
}

int emu_flags(int rank, int nn, int h, int kind, const uint64_t* keys, uint32_t n, const void* links,
              const uint32_t* metas, const double* u, double rthr, double cthr, int mn, int mx, int8_t* out)
{
    auto const* L = static_cast<const face_link*>(links);
#define X(R_, N_, H_)                                                          \
    if (rank == R_ && nn == N_ && h == H_)                                     \
    {                                                                          \
        flags<R_, N_, H_>(kind, keys, n, L, metas, u, rthr, cthr, mn, mx, out); \
        return 0;                                                              \
    }
    EMU_SHAPES(X)
#undef X
    return -1;
}

int emu_regrid(int rank, int nn, int depth, const uint64_t* keys, uint32_t n, const int8_t* fl, const void* links,
               const uint32_t* metas, const double* u, int max_level, uint64_t* keys_out, double* u_out, uint32_t cap,
               uint32_t* summary, uint32_t* maps)
{
    auto const* L = static_cast<const face_link*>(links);
#define X(R_, N_, H_)                                                                                       \
    if (rank == R_ && nn == N_)                                                                             \
        return regrid<R_, N_>(depth, keys, n, fl, L, metas, u, max_level, keys_out, u_out, cap, summary, maps);
    EMU_SHAPES(X)
#undef X
    return -1;
}

// Run list (head, len pairs in leaf order) of a valid table; returns the count.
int emu_runs(int rank, int depth, const uint64_t* keys, uint32_t n, const void* links, const uint32_t* metas,
             uint32_t K, uint32_t* out)
{
    auto const* L     = static_cast<const face_link*>(links);
    int         count = 0;
    for (uint32_t i = 0; i < n; ++i)
    {
        const bool h = rank == 2 ? run_head<2>(keys, depth, metas, i, K) : run_head<3>(keys, depth, metas, i, K);
        if (!h) continue;
        out[2 * count]     = i;
        out[2 * count + 1] = rank == 2 ? run_length<2>(keys, depth, L, metas, i, K) : run_length<3>(keys, depth, L, metas, i, K);
        ++count;
    }
    return count;
}

// dtdx: [17][3]; vel: 3 values (vp, vm and sgn are derived as on the device).
int emu_adj_step(int rank, int nn, const uint64_t* keys, uint32_t n, const void* links, const uint32_t* metas,
                 const double* lam_next, const double* u, const double* dtdx, const double* vel, double* lam,
                 double* vpart)
{
    auto const*    L = static_cast<const face_link*>(links);
    adjoint_params prm{};
    std::memcpy(prm.dtdx, dtdx, sizeof(prm.dtdx));
    for (int d = 0; d < 3; ++d)
    {
        prm.vp[d]  = 0.5 * (vel[d] + std::abs(vel[d]));
        prm.vm[d]  = 0.5 * (vel[d] - std::abs(vel[d]));
        prm.sgn[d] = vel[d] > 0.0 ? 1.0 : (vel[d] < 0.0 ? -1.0 : 0.0);
    }
#define X(R_, N_, H_)                                                              \
    if (rank == R_ && nn == N_)                                                    \
    {                                                                              \
        adj_step<R_, N_>(keys, n, L, metas, lam_next, u, prm, lam, vpart);         \
        return 0;                                                                  \
    }
    EMU_SHAPES(X)
#undef X
    return -1;
}

int emu_adj_regrid(int rank, int nn, const double* lam_new, const uint32_t* maps, uint32_t n_old, double* lam_old)
{
    std::vector<uint8_t> mark(maps, maps + n_old), head(maps + n_old, maps + 2 * n_old);
    const uint32_t*      counts = maps + 2 * n_old;
    const uint32_t*      pos    = maps + 3 * n_old;
#define X(R_, N_, H_)                                                          \
    if (rank == R_ && nn == N_)                                                \
    {                                                                          \
        adj_regrid<R_, N_>(lam_new, mark.data(), head.data(), counts, pos, lam_old, n_old); \
        return 0;                                                              \
    }
    EMU_SHAPES(X)
#undef X
    return -1;
}

// The 27 slots at offsets (dx, dy, dz) in {-1, 0, 1}^3 (index 9 (dz+1) +
// 3 (dy+1) + (dx+1)) of slot m in a single-level 3D tree at `level`, by the
// fused kernel's dilated-integer arithmetic.
int emu_uniform_neighbors(uint32_t level, uint32_t m, uint32_t* out)
{
    const morton_masks3 mk(level);
    int                 i = 0;
#define N3(DZ, DY, DX) out[i++] = mk.at<DX, DY, DZ>(m);
#define ROW(DZ, DY) N3(DZ, DY, -1) N3(DZ, DY, 0) N3(DZ, DY, 1)
#define PLANE(DZ) ROW(DZ, -1) ROW(DZ, 0) ROW(DZ, 1)
    PLANE(-1) PLANE(0) PLANE(1)
#undef PLANE
#undef ROW
#undef N3
    return 0;
}

// ---- any-shape (box) kernels: ext = extents of layout dims 0..rank-1 ------

int emu_step_box(int rank, const int* ext, const uint64_t* keys, uint32_t n, const void* links,
                 const uint32_t* metas, const double* u, double* un, const double* dtdx, const double* vel,
                 const double* hs)
{
    auto const* L = static_cast<const face_link*>(links);
    if (rank == 2)
        step_box<2>(ext, keys, n, L, metas, u, un, dtdx, vel, hs);
    else
        step_box<3>(ext, keys, n, L, metas, u, un, dtdx, vel, hs);
    return 0;
}

int emu_padded_box(int rank, const int* ext, int h, uint32_t n, const void* links, const uint32_t* metas,
                   const double* u, double* pad)
{
    auto const* L = static_cast<const face_link*>(links);
    if (rank == 2)
        padded_box<2>(ext, h, n, L, metas, u, pad);
    else
        padded_box<3>(ext, h, n, L, metas, u, pad);
    return 0;
}

int emu_flags_box(int rank, const int* ext, int h, int kind, const uint64_t* keys, uint32_t n, const void* links,
                  const uint32_t* metas, const double* u, double rthr, double cthr, int mn, int mx, int8_t* out)
{
    auto const* L = static_cast<const face_link*>(links);
    if (rank == 2)
        flags_box<2>(ext, h, kind, keys, n, L, metas, u, rthr, cthr, mn, mx, out);
    else
        flags_box<3>(ext, h, kind, keys, n, L, metas, u, rthr, cthr, mn, mx, out);
    return 0;
}

int emu_transfer_box(int rank, const int* ext, int depth, const uint64_t* keys, uint32_t n, const uint32_t* maps,
                     const double* u, uint64_t* keys_out, double* u_out)
{
    if (rank == 2)
        transfer_box<2>(ext, depth, keys, n, maps, u, keys_out, u_out);
    else
        transfer_box<3>(ext, depth, keys, n, maps, u, keys_out, u_out);
    return 0;
}

int emu_adj_step_box(int rank, const int* ext, const uint64_t* keys, uint32_t n, const void* links,
                     const uint32_t* metas, const double* lam_next, const double* u, const double* dtdx,
                     const double* vel, double* lam, double* vpart)
{
    auto const*          L   = static_cast<const face_link*>(links);
    const adjoint_params prm = make_adjoint_params(dtdx, vel);
    if (rank == 2)
        adj_step_box<2>(ext, keys, n, L, metas, lam_next, u, prm, lam, vpart);
    else
        adj_step_box<3>(ext, keys, n, L, metas, lam_next, u, prm, lam, vpart);
    return 0;
}

int emu_adj_regrid_box(int rank, const int* ext, const double* lam_new, const uint32_t* maps, uint32_t n_old,
                       double* lam_old)
{
    if (rank == 2)
        adj_regrid_box<2>(ext, lam_new, maps, n_old, lam_old);
    else
        adj_regrid_box<3>(ext, lam_new, maps, n_old, lam_old);
    return 0;
}

} // extern "C"
