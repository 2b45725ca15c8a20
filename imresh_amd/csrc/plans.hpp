// Compile-time line-FFT plans.  Each supported length has one radix sequence
// (forward order; the inverse runs it reversed).  The same list drives the
// kernel instantiations and the host-side twiddle tables, so both always agree.
//
// IMR_PLANS(X) entries: X(group, n, radices...).  `group` splits the kernel
// instantiations over several translation units (parallel compilation).
#pragma once

#include "common.hpp"

#define IMR_PLANS(X)                                                            \
    X(0, 1, 1) X(0, 2, 2) X(0, 3, 3) X(0, 4, 4) X(0, 5, 5) X(0, 8, 8)           \
    X(0, 12, 12) X(0, 16, 16) X(0, 32, 8, 4) X(0, 60, 4, 15) X(0, 64, 8, 8)     \
    X(1, 128, 16, 8) X(1, 256, 16, 16) X(1, 512, 8, 8, 8)                       \
    X(2, 1024, 16, 8, 8) X(2, 2048, 16, 16, 8)                                  \
    X(3, 4096, 16, 16, 16) X(3, 8192, 16, 8, 8, 8)                              \
    X(4, 96, 8, 12) X(4, 200, 8, 5, 5) X(4, 384, 16, 8, 3) X(4, 768, 16, 16, 3) \
    X(5, 1280, 16, 16, 5) X(5, 1536, 16, 16, 6) X(5, 2560, 16, 16, 10)          \
    X(5, 3072, 16, 16, 12)                                                      \
    X(6, 160, 16, 10) X(6, 192, 16, 12) X(6, 320, 16, 5, 4) X(6, 640, 16, 8, 5) \
    X(6, 5120, 16, 16, 5, 4) X(6, 6144, 16, 16, 8, 3)                           \
    X(7, 16384, 16, 16, 8, 8)

#define IMR_NUM_PLAN_GROUPS 8

namespace imr {

constexpr int kElemsPerThread = 16;      // E: complex values per thread and pass
constexpr int kRowThreadsTarget = 256;   // row workgroup size target
constexpr int kColMaxCols = 8;           // columns per column workgroup (64 B per row)
constexpr int kLdsBudget = 144 * 1024;   // dynamic LDS per workgroup (of 160 KiB)

// E: complex values per thread and pass (the thread count per line follows)
template <int E_, int N, int... Rs>
struct PlanE {
    static constexpr int n = N;
    static constexpr int np = sizeof...(Rs);
    static constexpr int rf[np] = {Rs...};
    static constexpr int E = E_;

    __host__ __device__ static constexpr int rad(bool inv, int i) { return inv ? rf[np - 1 - i] : rf[i]; }
    __host__ __device__ static constexpr int ns(bool inv, int i)
    {
        int s = 1;
        for (int q = 0; q < i; ++q) s *= rad(inv, q);
        return s;
    }
    // twiddle table: [forward sequence: n-1 entries][inverse sequence: n-1 entries]
    __host__ __device__ static constexpr int twoff(bool inv, int i)
    {
        int o = inv ? N - 1 : 0;
        for (int q = 0; q < i; ++q) o += ns(inv, q) * (rad(inv, q) - 1);
        return o;
    }
    __host__ __device__ static constexpr int calc_T()
    {
        int T = 1;
        for (int i = 0; i < np; ++i) {
            const int R = rf[i];
            if (R == 1) continue;
            const int nb = N / R, per = E / R;
            const int t = (nb + per - 1) / per;
            T = t > T ? t : T;
        }
        return T;
    }
    static constexpr int T = calc_T();
    // rows per row workgroup
    static constexpr int L = T >= kRowThreadsTarget ? 1 : kRowThreadsTarget / T;
    __host__ __device__ static constexpr int calc_C()
    {
        int C = kColMaxCols;
        while (C > 1 && (C * T > 1024 || (long) sizeof(float) * 2 * lds_elems_c(C * N) > kLdsBudget))
            C /= 2;
        return C;
    }
    __host__ __device__ static constexpr int lds_elems_c(int m) { return m + (m >> 5) + 1; }
    static constexpr int C = calc_C();
    static_assert(L * T <= 1024, "row workgroup too large");
    static_assert(C * T <= 1024, "column workgroup too large");
};

template <int N, int... Rs>
using Plan = PlanE<kElemsPerThread, N, Rs...>;

// host view of a plan
struct PlanInfo {
    int group, n, np, T, L, C;
    int rad[kMaxPasses];
};

}  // namespace imr
