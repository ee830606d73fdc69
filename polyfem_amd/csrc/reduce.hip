// Deterministic reductions and small utility kernels.
#include "dispatch.hpp"

namespace pfgpu {

namespace {

// in[i0], in[i0 + BLOCK], ... below i1, added in that order with eight loads
// in flight (the order, and so the result, is that of the plain loop)
__device__ __forceinline__ double strided_sum(const double *__restrict__ in, int i0, int i1)
{
    double v = 0;
    int i = i0;
    for (; i + 7 * BLOCK < i1; i += 8 * BLOCK) {
        double a[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = in[i + k * BLOCK];
#pragma unroll
        for (int k = 0; k < 8; ++k) v += a[k];
    }
    for (; i < i1; i += BLOCK) v += in[i];
    return v;
}

// Single block: thread t sums in[t], in[t+256], ... in order, then a fixed
// tree.  The summation order depends only on n, never on scheduling.
__global__ __launch_bounds__(BLOCK) void k_sum_doubles(const double *__restrict__ in, int n, double *__restrict__ out)
{
    __shared__ double s[BLOCK];
    const int t = threadIdx.x;
    const double v = strided_sum(in, t, n);
    s[t] = v;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) out[0] = s[0];
}

// First level for long inputs: block b sums the contiguous range
// [b seg, (b+1) seg) (thread-strided, then a fixed tree) into part[b].
__global__ __launch_bounds__(BLOCK) void k_sum_doubles_blocks(const double *__restrict__ in, int n, int seg,
                                                              double *__restrict__ part)
{
    __shared__ double s[BLOCK];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * seg, i1 = min(n, i0 + seg);
    const double v = strided_sum(in, i0 + t, i1);
    s[t] = v;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = s[0];
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_count_nonfinite(const T *__restrict__ v, int64_t n, int *__restrict__ flag)
{
    bool bad = false;
    for (int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += int64_t(gridDim.x) * BLOCK)
        bad |= !isfinite(v[i]);
    // a plain store of the same value from any thread: no read-modify-write
    if (bad) flag[0] = 1;
}

// scalar CSR column indices from the block pattern: row 3i+m lists 3j+n
__global__ __launch_bounds__(BLOCK) void k_expand_inner(int n_nodes, const int32_t *__restrict__ brow_ptr,
                                                        const int32_t *__restrict__ bcol, int32_t *__restrict__ inner)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes) return;
    const int r0 = brow_ptr[i], deg = brow_ptr[i + 1] - r0;
    for (int m = 0; m < 3; ++m)
        for (int p = 0; p < deg; ++p)
            for (int n = 0; n < 3; ++n) inner[9 * int64_t(r0) + int64_t(m) * 3 * deg + 3 * p + n] = 3 * bcol[r0 + p] + n;
}

} // namespace

void sum_doubles_into(Handle &h, const double *in, int n, double *out, double *part)
{
    // one block for short inputs; above that a fixed grid of SUM_BLOCKS
    // contiguous ranges, then one block over their sums.  The order depends
    // on n only.
    if (n <= 32 * BLOCK) {
        k_sum_doubles<<<1, BLOCK, 0, h.stream>>>(in, n, out);
    } else {
        const int seg = (n + SUM_BLOCKS - 1) / SUM_BLOCKS;
        k_sum_doubles_blocks<<<SUM_BLOCKS, BLOCK, 0, h.stream>>>(in, n, seg, part);
        k_sum_doubles<<<1, BLOCK, 0, h.stream>>>(part, SUM_BLOCKS, out);
    }
    PF_LAUNCH_CHECK();
}

void sum_doubles(Handle &h, const double *in, int n, double *out)
{
    if (n > 32 * BLOCK && h.sum_part.bytes() < sizeof(double) * SUM_BLOCKS) h.sum_part.reserve(sizeof(double) * SUM_BLOCKS);
    sum_doubles_into(h, in, n, out, h.sum_part.as<double>());
}

bool all_finite(Handle &h, const void *dev, int64_t n, bool is_float)
{
    h.red_out.reserve(sizeof(int) * 4);
    int *flag = h.red_out.as<int>();
    PF_HIP(hipMemsetAsync(flag, 0, sizeof(int), h.stream));
    const unsigned g = unsigned(std::min<int64_t>(4096, (n + BLOCK - 1) / BLOCK + 1));
    if (is_float)
        k_count_nonfinite<float><<<g, BLOCK, 0, h.stream>>>(static_cast<const float *>(dev), n, flag);
    else
        k_count_nonfinite<double><<<g, BLOCK, 0, h.stream>>>(static_cast<const double *>(dev), n, flag);
    PF_LAUNCH_CHECK();
    int host = 0;
    PF_HIP(hipMemcpyAsync(&host, flag, sizeof(int), hipMemcpyDeviceToHost, h.stream));
    PF_HIP(hipStreamSynchronize(h.stream));
    return host == 0;
}

void build_device_pattern(Handle &h)
{
    if (h.d_outer.get()) return;
    h.d_outer = upload(h.pat.outer, h.stream);
    DevBuf bcol = upload(h.pat.bcol, h.stream);
    DevBuf brow = upload(h.pat.brow_ptr, h.stream);
    h.d_inner.alloc(sizeof(int32_t) * size_t(h.pat.nnz));
    k_expand_inner<<<grid_for(h.n_nodes), BLOCK, 0, h.stream>>>(h.n_nodes, brow.as<int32_t>(), bcol.as<int32_t>(),
                                                                 h.d_inner.as<int32_t>());
    PF_LAUNCH_CHECK();
    PF_HIP(hipStreamSynchronize(h.stream));
}

} // namespace pfgpu
