// Outer steps of the four-step column transform of length H = 8 * H2
// (used for H = 8192, whose own column tile holds only two columns).
// With n = H2*n1 + n2 and k = k1 + 8*k2,
//   X[k1 + 8 k2] = sum_n2 w_H2^(n2 k2) [ w_H^(n2 k1) sum_n1 x[H2 n1 + n2] w_8^(n1 k1) ],
// so the forward step runs 8-point DFTs down each column at stride H2 rows
// and multiplies by w_H^(n2 k1) in place (block k1 of H2 rows then holds
// the inputs of the length-H2 transforms of frequencies k1 + 8 k2), and the
// inverse step undoes it with conjugate twiddles.  Every access moves whole
// rows (64 columns per wave: 512-byte segments), 16 bytes per element.
#include "internal.hpp"
#include "fft_device.hpp"

namespace imr {

namespace {

template <int DIR>
__global__ void __launch_bounds__(256) k_step8(Step8Args a)
{
    const int b = blockIdx.z;
    if (a.done != nullptr && a.done[b * kFlagStride]) return;
    const int c = (int) blockIdx.x * 64 + (int) (threadIdx.x & 63);
    const int n2 = (int) blockIdx.y * 4 + (int) (threadIdx.x >> 6);
    if (c >= a.W || n2 >= a.H2) return;
    cf *X = (cf *) a.X + (size_t) b * a.istride;
    const cf *tw = (const cf *) a.tw;
    const size_t rs = (size_t) a.H2 * a.W;   // between the 8 rows of a DFT
    const size_t base = (size_t) n2 * a.W + c;
    cf v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = X[base + r * rs];
    if constexpr (DIR < 0) {
        bf8<-1>(v);
#pragma unroll
        for (int k = 1; k < 8; ++k) v[k] = cmul(v[k], tw[n2 * k]);
    } else {
#pragma unroll
        for (int k = 1; k < 8; ++k) v[k] = cmulc(v[k], tw[n2 * k]);
        bf8<+1>(v);
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) X[base + r * rs] = v[r];
}

}  // namespace

int launch_step8(int dir, const Step8Args &a, int batch, hipStream_t s)
{
    if (a.H2 % 4 != 0) return IMRESH_ERR_INVALID;
    const dim3 grid((a.W + 63) / 64, a.H2 / 4, batch > 1 ? batch : 1), block(256);
    if (dir < 0) IMR_LAUNCH((k_step8<-1>), grid, block, 0, s, a);
    else IMR_LAUNCH((k_step8<+1>), grid, block, 0, s, a);
    return imr::launch_status(__FILE__, __LINE__);
}

}  // namespace imr
