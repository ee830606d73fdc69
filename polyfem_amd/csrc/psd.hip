// Projection of every element's local Hessian onto the positive semidefinite
// cone: the reference's project_to_psd option (NLAssembler::assemble_hessian,
// Assembler.cpp:645-771, calling ipc::project_to_psd of ipc-toolkit@b40e9c0
// with its CLAMP method): with A = V diag(d) V^T, eigenvalues <= 0 become 0;
// a matrix whose eigenvalues are all >= 0 is left as it is.
//
// One wavefront per element.  The symmetric 3n x 3n matrix (n = 4 or 10
// nodes) is read from the element-block tiles k_hessian writes.  A Cholesky
// factorisation of A + delta I (delta = psd_shift x the largest diagonal
// entry) first tests whether every eigenvalue exceeds -delta; such a matrix
// is left as it is (the exact projection would move it by at most delta).
// Otherwise A is diagonalised in LDS by the cyclic Jacobi method with the
// round-robin ordering (n*3/2 disjoint rotations per round, all applied at
// once, n*3-1 rounds per sweep) until the off-diagonal part is at rounding
// level, and written back when some eigenvalue is negative.
//
// FP32 matrices take a second method built on 32x32 matrix products on
// the matrix cores: A_+ = (A + A sign(A)) / 2, the matrix sign function from
// Newton-Schulz iterations on X0 = A / |A|_F (every eigenvalue in [-1, 1]):
// NS_FAST steps of the quintic X <- X (a I + b X^2 + c X^4), which lifts small
// eigenvalues by about 3.4x per step into [0.7, 1.2] in magnitude (the
// coefficients of K. Jordan's Muon orthogonalisation, 2024), then NS_POLISH
// steps of the cubic X <- X (3 I - X^2) / 2, which converges quadratically to
// sign(A) there.  An eigenvalue too small to be lifted in time contributes
// at most its own magnitude, below 1e-6 of |A| after these steps: the result
// agrees with the eigendecomposition's to FP32 rounding of |A|.
#include "dispatch.hpp"

namespace pfgpu {

namespace {

// round-robin pairing of N indices: round r pairs pos[k] with pos[N-1-k]
template <int N> struct JacobiRounds {
    int8_t p[(N - 1) * (N / 2)], q[(N - 1) * (N / 2)];
    constexpr JacobiRounds() : p(), q()
    {
        int pos[N] = {};
        for (int i = 0; i < N; ++i) pos[i] = i;
        for (int r = 0; r < N - 1; ++r) {
            for (int k = 0; k < N / 2; ++k) {
                const int a = pos[k], b = pos[N - 1 - k];
                p[r * (N / 2) + k] = int8_t(a < b ? a : b);
                q[r * (N / 2) + k] = int8_t(a < b ? b : a);
            }
            const int last = pos[N - 1];
            for (int i = N - 1; i > 1; --i) pos[i] = pos[i - 1];
            pos[1] = last;
        }
    }
};
__constant__ JacobiRounds<12> c_jr12 = JacobiRounds<12>();
__constant__ JacobiRounds<30> c_jr30 = JacobiRounds<30>();

template <int N> __device__ __forceinline__ const JacobiRounds<N> &rounds()
{
    if constexpr (N == 12) return c_jr12;
    else return c_jr30;
}

// element-local node pair of upper block s (a <= b), in k_hessian's order
template <int NB> struct BlockPairs {
    int8_t a[NB * (NB + 1) / 2], b[NB * (NB + 1) / 2];
    constexpr BlockPairs() : a(), b()
    {
        int s = 0;
        for (int i = 0; i < NB; ++i)
            for (int j = i; j < NB; ++j) {
                a[s] = int8_t(i);
                b[s] = int8_t(j);
                ++s;
            }
    }
};
__constant__ BlockPairs<4> c_bp4 = BlockPairs<4>();
__constant__ BlockPairs<10> c_bp10 = BlockPairs<10>();

template <int NB> __device__ __forceinline__ const BlockPairs<NB> &block_pairs()
{
    if constexpr (NB == 4) return c_bp4;
    else return c_bp10;
}

constexpr int PSD_WAVE = 64;

// relative shift of the positive-definiteness test: far below the parity
// tolerances (1e-10 in FP64, 1e-4 for FP32 values) and far above the
// rounding of a Cholesky factorisation of order 30
template <typename T> constexpr T psd_shift() { return sizeof(T) == 8 ? T(1e-12) : T(2e-6); }

template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = PSD_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, PSD_WAVE);
    return v;
}

template <typename T, int ORDER>
__global__ __launch_bounds__(PSD_WAVE) void k_psd_project(int n_elements, T *__restrict__ hloc)
{
    constexpr int NB = ORDER == 1 ? 4 : 10, N = 3 * NB, NP = N / 2, NBLK = NB * (NB + 1) / 2;
    constexpr int MAX_SWEEPS = 20;
    // squared relative off-diagonal size at which the iteration stops
    // (the rounding floor of the off-diagonal part is about (N eps)^2 of the diagonal)
    constexpr double TOL2 = sizeof(T) == 8 ? 1e-26 : 1e-10;
    __shared__ T A[N][N + 1], V[N][N + 1], CS[NP][2];
    __shared__ int PQ[NP][2];
    const int e = blockIdx.x, lane = threadIdx.x;
    if (e >= n_elements) return; // uniform over the workgroup
    // tile layout [tile][s][element in tile][9]: 64 elements per tile
    T *base = hloc + (int64_t(e) / 64) * NBLK * 576 + (e % 64) * 9;
    const BlockPairs<NB> &bp = block_pairs<NB>();
    for (int i = lane; i < NBLK * 9; i += PSD_WAVE) {
        const int s = i / 9, j = i % 9, a = bp.a[s], b = bp.b[s];
        const T v = base[int64_t(s) * 576 + j];
        A[3 * a + j / 3][3 * b + j % 3] = v;
        A[3 * b + j % 3][3 * a + j / 3] = v;
    }
    __syncthreads();
    // positive-definiteness test: Cholesky of A + delta I in V (lower triangle)
    T dmax = T(0);
    for (int k = lane; k < N; k += PSD_WAVE) dmax = std::fmax(dmax, std::fabs(A[k][k]));
#pragma unroll
    for (int o = PSD_WAVE / 2; o > 0; o >>= 1) dmax = std::fmax(dmax, __shfl_xor(dmax, o, PSD_WAVE));
    const T delta = psd_shift<T>() * dmax;
    for (int i = lane; i < N * N; i += PSD_WAVE) V[i / N][i % N] = A[i / N][i % N] + (i / N == i % N ? delta : T(0));
    __syncthreads();
    bool pd = true;
    for (int k = 0; k < N; ++k) {
        const T d = V[k][k]; // the same value in every lane: a uniform decision
        if (!(d > T(0))) {
            pd = false;
            break;
        }
        const T inv = T(1) / std::sqrt(d);
        __syncthreads();
        for (int i = k + 1 + lane; i < N; i += PSD_WAVE) V[i][k] *= inv;
        __syncthreads();
        // trailing lower triangle: V[i][j] -= V[i][k] V[j][k], k < j <= i
        constexpr int NN = N * N;
        for (int idx = lane; idx < NN; idx += PSD_WAVE) {
            const int i = idx / N, j = idx % N;
            if (j > k && j <= i) V[i][j] -= V[i][k] * V[j][k];
        }
        __syncthreads();
    }
    if (pd) return; // every eigenvalue above -delta: A stays
    for (int i = lane; i < N * N; i += PSD_WAVE) V[i / N][i % N] = i / N == i % N ? T(1) : T(0);
    __syncthreads();
    const JacobiRounds<N> &jr = rounds<N>();
    // this lane's (pair, column) items of the row and column passes: item
    // i = lane + 64 it is pair i / N, column i % N (the same in every round)
    constexpr int ITEMS = NP * N, NIT = (ITEMS + PSD_WAVE - 1) / PSD_WAVE;
    int ik[NIT], ij[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int i = lane + PSD_WAVE * it;
        ik[it] = i < ITEMS ? i / N : -1;
        ij[it] = i < ITEMS ? i % N : 0;
    }
    for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
        T off = T(0), dia = T(0);
        for (int i = lane; i < N * N; i += PSD_WAVE) {
            const T v = A[i / N][i % N];
            if (i / N == i % N)
                dia += v * v;
            else
                off += v * v;
        }
        off = wave_sum(off);
        dia = wave_sum(dia);
        // the same values in every lane: a uniform decision
        if (!(off > T(TOL2) * dia)) break;
        for (int r = 0; r < N - 1; ++r) {
            if (lane < NP) {
                const int p = jr.p[r * NP + lane], q = jr.q[r * NP + lane];
                PQ[lane][0] = p;
                PQ[lane][1] = q;
                const T apq = A[p][q];
                T c = T(1), s = T(0);
                if (apq != T(0)) {
                    // the rotation that zeroes A[p][q] (the classical Jacobi step)
                    const T th = (A[q][q] - A[p][p]) / (T(2) * apq);
                    const T t = (th >= T(0) ? T(1) : T(-1)) / (std::fabs(th) + std::sqrt(th * th + T(1)));
                    c = T(1) / std::sqrt(t * t + T(1));
                    s = t * c;
                }
                CS[lane][0] = c;
                CS[lane][1] = s;
            }
            __syncthreads();
            // rows p, q of every pair (the pairs are disjoint), and the
            // eigenvector matrix's columns p, q
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int k = ik[it], j = ij[it];
                if (k >= 0) {
                    const int p = PQ[k][0], q = PQ[k][1];
                    const T c = CS[k][0], s = CS[k][1], ap = A[p][j], aq = A[q][j];
                    A[p][j] = c * ap - s * aq;
                    A[q][j] = s * ap + c * aq;
                    const T vp = V[j][p], vq = V[j][q];
                    V[j][p] = c * vp - s * vq;
                    V[j][q] = s * vp + c * vq;
                }
            }
            __syncthreads();
            // columns p, q; the entry the rotation annihilates is set to zero
            // rather than left at rounding level
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int k = ik[it], j = ij[it];
                if (k >= 0) {
                    const int p = PQ[k][0], q = PQ[k][1];
                    const T c = CS[k][0], s = CS[k][1];
                    const T ap = A[j][p], aq = A[j][q];
                    A[j][p] = j == q ? T(0) : c * ap - s * aq;
                    A[j][q] = j == p ? T(0) : s * ap + c * aq;
                }
            }
            __syncthreads();
        }
    }
    // eigenvalues on the diagonal; nothing to do when none is negative
    bool neg = false;
    for (int k = lane; k < N; k += PSD_WAVE) neg |= A[k][k] < T(0);
    if (!__any(neg)) return;
    for (int i = lane; i < NBLK * 9; i += PSD_WAVE) {
        const int s = i / 9, j = i % 9, r = 3 * bp.a[s] + j / 3, c = 3 * bp.b[s] + j % 3;
        T acc = T(0);
#pragma unroll 6
        for (int k = 0; k < N; ++k) {
            const T d = A[k][k];
            if (d > T(0)) acc += V[r][k] * d * V[c][k];
        }
        base[int64_t(s) * 576 + j] = acc;
    }
}

// ---- Newton-Schulz projection on the matrix cores (FP32, P2)

typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int NS_FAST = 8, NS_POLISH = 10;
constexpr float NS_A = 3.4445f, NS_B = -4.7750f, NS_C = 2.0315f;
constexpr int NSS = 33; // row stride of a 32x32 matrix in LDS (conflict-free rows and columns)

// P Q for 32x32 matrices in LDS, with v_mfma_f32_32x32x2f32 over k = 2s, 2s+1:
// lane l supplies P[l % 32][2s + l / 32] and Q[2s + l / 32][l % 32]
__device__ __forceinline__ floatx16 ns_mul(const float *P, const float *Q, int i0, int h)
{
    floatx16 acc = {};
#pragma unroll
    for (int s = 0; s < 16; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(P[i0 * NSS + 2 * s + h], Q[(2 * s + h) * NSS + i0], acc, 0, 0, 0);
    return acc;
}
// row of accumulator r of lane half h (the column is l % 32)
__device__ __forceinline__ int ns_row(int r, int h) { return 8 * (r / 4) + 4 * h + (r % 4); }
__device__ __forceinline__ void ns_store(float *D, const floatx16 &v, int i0, int h)
{
#pragma unroll
    for (int r = 0; r < 16; ++r) D[ns_row(r, h) * NSS + i0] = v[r];
}

// One wavefront per P2 element (30x30 padded to 32x32 with zeros, which
// every polynomial in X keeps zero) or per two P1 elements (12x12 matrices
// at rows and columns 0 and 16 of a block-diagonal 32x32, each scaled by its
// own norm: polynomials in a block-diagonal X act on each block alone).  The
// workgroup is the wavefront, so the barriers only order the LDS accesses of
// its lanes.
template <int ORDER> __global__ __launch_bounds__(64) void k_psd_ns(int n_elements, float *__restrict__ hloc)
{
    constexpr int NB = ORDER == 1 ? 4 : 10, N = 3 * NB, NBLK = NB * (NB + 1) / 2, EPW = ORDER == 1 ? 2 : 1;
    // A, the iterate X and a temporary W (X^2, then the polynomial factor):
    // 12.7 KB, 12 wavefronts per CU
    __shared__ float A[32 * NSS], X[32 * NSS], W[32 * NSS];
    const int lane = threadIdx.x, i0 = lane % 32, h = lane / 32;
    const int e0 = int(blockIdx.x) * EPW;
    if (e0 >= n_elements) return; // uniform over the workgroup
    const int ne = min(EPW, n_elements - e0);
    const BlockPairs<NB> &bp = block_pairs<NB>();
    auto base = [&](int e) { return hloc + (int64_t(e) / 64) * NBLK * 576 + (e % 64) * 9; };
    for (int k = lane; k < 32 * NSS; k += 64) A[k] = 0.f;
    __syncthreads();
    for (int i = lane; i < ne * NBLK * 9; i += 64) {
        const int k = i / (NBLK * 9), ii = i - k * NBLK * 9, s = ii / 9, j = ii % 9;
        const int r = 16 * k + 3 * bp.a[s] + j / 3, c = 16 * k + 3 * bp.b[s] + j % 3;
        const float v = base(e0 + k)[int64_t(s) * 576 + j];
        A[r * NSS + c] = v;
        A[c * NSS + r] = v;
    }
    __syncthreads();
    // the Frobenius norm of each element's matrix
    float sc[EPW];
#pragma unroll
    for (int k = 0; k < EPW; ++k) {
        float ss = 0.f;
        for (int q = lane; q < N * N; q += 64) {
            const float v = A[(16 * k + q / N) * NSS + 16 * k + q % N];
            ss = fmaf(v, v, ss);
        }
        ss = wave_sum(ss);
        sc[k] = ss > 0.f ? 1.f / std::sqrt(ss) : 0.f; // a zero matrix stays zero
    }
    for (int k = lane; k < 32 * NSS; k += 64) X[k] = A[k] * (EPW == 1 || k / NSS < 16 ? sc[0] : sc[EPW - 1]);
    __syncthreads();
    for (int it = 0; it < NS_FAST + NS_POLISH; ++it) {
        const floatx16 x2 = ns_mul(X, X, i0, h);
        floatx16 m;
        if (it < NS_FAST) {
            ns_store(W, x2, i0, h);
            __syncthreads();
            const floatx16 x4 = ns_mul(W, W, i0, h);
            __syncthreads(); // every lane has read W
#pragma unroll
            for (int r = 0; r < 16; ++r) m[r] = fmaf(NS_C, x4[r], NS_B * x2[r]) + (ns_row(r, h) == i0 ? NS_A : 0.f);
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) m[r] = -0.5f * x2[r] + (ns_row(r, h) == i0 ? 1.5f : 0.f);
        }
        ns_store(W, m, i0, h);
        __syncthreads();
        const floatx16 xn = ns_mul(X, W, i0, h);
        __syncthreads(); // every lane has read X
        ns_store(X, xn, i0, h);
        __syncthreads();
    }
    const floatx16 ax = ns_mul(A, X, i0, h);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = ns_row(r, h);
        W[i * NSS + i0] = 0.5f * (A[i * NSS + i0] + ax[r]);
    }
    __syncthreads();
    for (int i = lane; i < ne * NBLK * 9; i += 64) {
        const int k = i / (NBLK * 9), ii = i - k * NBLK * 9, s = ii / 9, j = ii % 9;
        const int r = 16 * k + 3 * bp.a[s] + j / 3, c = 16 * k + 3 * bp.b[s] + j % 3;
        base(e0 + k)[int64_t(s) * 576 + j] = W[r * NSS + c];
    }
}

} // namespace

// hloc: the element-block tiles of k_hessian for every element of the handle
void psd_project_blocks(Handle &h, void *hloc)
{
    if (h.n_elements == 0) return;
    auto go = [&]<typename T>() {
        if constexpr (sizeof(T) == 4) {
            if (h.order == 1)
                k_psd_ns<1><<<(h.n_elements + 1) / 2, 64, 0, h.stream>>>(h.n_elements, static_cast<float *>(hloc));
            else
                k_psd_ns<2><<<h.n_elements, 64, 0, h.stream>>>(h.n_elements, static_cast<float *>(hloc));
        } else {
            if (h.order == 1)
                k_psd_project<T, 1><<<h.n_elements, PSD_WAVE, 0, h.stream>>>(h.n_elements, static_cast<T *>(hloc));
            else
                k_psd_project<T, 2><<<h.n_elements, PSD_WAVE, 0, h.stream>>>(h.n_elements, static_cast<T *>(hloc));
        }
    };
    if (h.precision == PFGPU_MIXED)
        go.template operator()<float>();
    else
        go.template operator()<double>();
    PF_LAUNCH_CHECK();
}

} // namespace pfgpu
