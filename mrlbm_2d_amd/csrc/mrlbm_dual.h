// mrlbm_dual.h -- forward-mode dual numbers (value + one directional derivative).
//
// The adjoint obtains the local Jacobian of each curved-wall and edge node by running the
// step's own per-node BC code (wall_bc_rotated, wall_bc_unrotated, edge_bc, collide) on
// Dual<T> with one input seeded per thread, so the derivative always follows the exact
// arithmetic of the forward step.
#pragma once
#include <hip/hip_runtime.h>

#include "mrlbm_lattice.h"

namespace mrlbm {

template <class B> struct Dual {
    B v, d;
    __host__ __device__ Dual() : v(0), d(0) {}
    __host__ __device__ Dual(B x) : v(x), d(0) {}
    __host__ __device__ Dual(B x, B dx) : v(x), d(dx) {}
    friend __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
    friend __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
    friend __device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
    friend __device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
    friend __device__ __forceinline__ Dual operator/(Dual a, Dual b)
    {
        const B q = a.v / b.v;
        return Dual(q, (a.d - q * b.d) / b.v);
    }
};

// Lattice constants of a dual type are those of its base type.
template <class B> struct Lat<Dual<B>> : Lat<B> {};

} // namespace mrlbm
