// ref_physics.cpp -- test infrastructure: the reference's own physics
// functions, compiled from its headers where they lie (include/solver/
// EulerPhysics.hpp, include/solver/AdvectionPhysics.hpp; built by
// `make -C oracle ref` into oracle/_ref/, which is not committed; the tree
// copied to a GPU host carries it, but only the CPU tests of
// tests/test_oracle_ref.py open it), behind a C face so that the tests can
// compare the oracle's restatement with them bit for bit.
// Both headers compile as they are with g++ -std=c++20; no stand-in is
// supplied for anything.
#include "solver/AdvectionPhysics.hpp"
#include "solver/EulerPhysics.hpp"

#include <cstddef>
#include <tuple>
#include <utility>

namespace
{

template <int R>
using vec = amr::containers::static_vector<double, R + 2>;

template <int R>
vec<R> load(const double* p)
{
    vec<R> v;
    for (int k = 0; k < R + 2; ++k) v[k] = p[k];
    return v;
}

template <int R>
void prim_to_cons(const double* prim, double* cons, double gamma)
{
    vec<R> c;
    EulerPhysics<R>::primitiveToConservative(load<R>(prim), c, gamma);
    for (int k = 0; k < R + 2; ++k) cons[k] = c[k];
}

template <int R>
void flux(const double* UL, const double* UR, double* F, int dir, double gamma)
{
    vec<R> f;
    EulerPhysics<R>::rusanovFlux(load<R>(UL), load<R>(UR), f, dir, gamma);
    for (int k = 0; k < R + 2; ++k) F[k] = f[k];
}

// getMaxSpeed reads field v of cell idx as std::get<v>(patches)[idx]: one
// single-cell "patch" per field.
template <int R, std::size_t... Is>
double max_speed(const double* U, int dir, double gamma, std::index_sequence<Is...>)
{
    const auto patches = std::make_tuple((U + Is)...);
    return EulerPhysics<R>::getMaxSpeed(patches, 0, dir, gamma);
}

} // namespace

extern "C"
{

int ref_euler_prim_to_cons(int rank, const double* prim, double* cons, double gamma)
{
    if (rank == 2) prim_to_cons<2>(prim, cons, gamma);
    else if (rank == 3) prim_to_cons<3>(prim, cons, gamma);
    else return 1;
    return 0;
}

int ref_euler_flux(int rank, const double* UL, const double* UR, double* F, int dir, double gamma)
{
    if (rank == 2) flux<2>(UL, UR, F, dir, gamma);
    else if (rank == 3) flux<3>(UL, UR, F, dir, gamma);
    else return 1;
    return 0;
}

double ref_euler_max_speed(int rank, const double* U, int dir, double gamma)
{
    if (rank == 2) return max_speed<2>(U, dir, gamma, std::make_index_sequence<4>{});
    return max_speed<3>(U, dir, gamma, std::make_index_sequence<5>{});
}

// AdvectionPhysics<rank>::rusanovFlux with the reference's fixed velocity
// (1, 0.5, 0) along direction dir.
double ref_advection_flux(int rank, double ul, double ur, int dir)
{
    amr::containers::static_vector<double, 1> L, Rv, F;
    L[0]  = ul;
    Rv[0] = ur;
    if (rank == 2) AdvectionPhysics<2>::rusanovFlux(L, Rv, F, dir, 1.4);
    else AdvectionPhysics<3>::rusanovFlux(L, Rv, F, dir, 1.4);
    return F[0];
}

} // extern "C"
