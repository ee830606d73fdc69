// Internal declarations shared by the host code and the kernels of libpfgpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pfgpu.h"

namespace pfgpu {

// Errors inside the library are C++ exceptions; every C entry point catches
// them and converts them to a status code plus pfgpu_last_error().
struct Error : std::runtime_error {
    pfgpu_status status;
    Error(pfgpu_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(pfgpu_status s, const std::string &m) { throw Error(s, m); }

inline void hip_check(hipError_t e, const char *what, const char *file, int line)
{
    if (e != hipSuccess) {
        const pfgpu_status s = (e == hipErrorOutOfMemory) ? PFGPU_ERR_OUT_OF_MEMORY : PFGPU_ERR_HIP;
        fail(s, std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")");
    }
}
#define PF_HIP(call) ::pfgpu::hip_check((call), #call, __FILE__, __LINE__)
#define PF_LAUNCH_CHECK() ::pfgpu::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Owning device allocation.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t bytes)
    {
        release();
        if (bytes) PF_HIP(hipMalloc(&p_, bytes));
        n_ = bytes;
    }
    // grow-only scratch
    void reserve(size_t bytes)
    {
        if (bytes > n_) alloc(bytes);
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p_); }
    void *get() const { return p_; }
    size_t bytes() const { return n_; }

  private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

// Chunked gather of the Hessian's pattern blocks for workgroups of `threads`
// threads: chunk starts (n_chunks + 1 blocks), threads contribution slots per
// chunk, per-block metadata (the kernels document the layout).
struct GatherPlan {
    DevBuf chunk, slot, meta, dst;
    int n_chunks = 0, threads = 0;
    unsigned grid = 0; // launch grid of a persistent gather over this plan (0: not yet known)
};

// Row-unit plan of the factored gather for workgroups of `threads` threads:
// chunks of whole node rows (per chunk boundary its first row, first unit
// and first pattern block, 4 ints, n_chunks + 1 entries), per unit (node
// incidence, in node_inc order) the stage positions of its n_local blocks,
// per pattern block the first stage position of its contributions within
// its chunk.
struct UnitPlan {
    DevBuf chunk_row, udst, bq0, lanes;
    int n_chunks = 0, threads = 0, split = 1, n_words = 0;
    unsigned grid = 0, grid_grad = 0;
};

// v into dst: in place when the size matches (device pointers held elsewhere,
// e.g. by captured graphs, stay valid), else into a new allocation
template <typename T> void upload_into(DevBuf &dst, const std::vector<T> &v, hipStream_t s);

template <typename T> DevBuf upload(const std::vector<T> &v, hipStream_t s)
{
    DevBuf b(v.size() * sizeof(T));
    if (!v.empty()) {
        PF_HIP(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        PF_HIP(hipStreamSynchronize(s));
    }
    return b;
}

template <typename T> void upload_into(DevBuf &dst, const std::vector<T> &v, hipStream_t s)
{
    if (dst.get() && dst.bytes() == v.size() * sizeof(T)) {
        PF_HIP(hipStreamSynchronize(s));
        PF_HIP(hipMemcpyAsync(dst.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        PF_HIP(hipStreamSynchronize(s));
    } else {
        dst = upload(v, s);
    }
}

// Problems of a batch whose element work is skipped (their CG has finished):
// flag[s] != 0 skips problem s; elem_seg / node_seg map elements / nodes to
// problems.  Empty (flag null) outside a batched solve.
struct SegSkip {
    const int32_t *elem_seg = nullptr, *node_seg = nullptr;
    const double *flag = nullptr;
    __host__ __device__ bool element(int e) const { return flag && flag[elem_seg[e]] != 0.0; }
    __host__ __device__ bool node(int i) const { return flag && flag[node_seg[i]] != 0.0; }
};

// Element data as the kernels see it.  Every per-element array is SoA
// (component-major, element-minor) so that one wavefront reading 64
// consecutive elements issues 64 consecutive addresses.
template <typename T> struct DevMesh {
    int n_elements;
    int n_nodes;
    int n_vertices;
    const int32_t *conn;  // [n_local][E]   dof-node ids
    const int32_t *vconn; // [4][E]         geometric vertex ids
    const T *geom;        // [10][E]        J^{-T} row-major (0..8), det J (9)
    const T *lame;        // [n_params][E]  lambda, mu (MooneyRivlin: c1, c2, k)
    int n_params;         // 2, or 3 for MooneyRivlin
};

// Sparsity pattern and the maps that scatter element-local results into it.
struct Pattern {
    int n_nodes = 0;
    int64_t n_blocks = 0;             // 3x3 node-pair blocks
    std::vector<int32_t> brow_ptr;    // [N+1]
    std::vector<int32_t> bcol;        // [n_blocks]
    std::vector<int32_t> outer;       // [3N+1] scalar CSC / CSR
    std::vector<int32_t> inner;       // [nnz]
    int64_t nnz = 0;
    // node -> (element, local node) incidence, sorted by element
    std::vector<int32_t> node_ptr;    // [N+1]
    std::vector<int32_t> node_inc;    // e * n_local + a
    // vertex -> (element, local vertex) incidence
    std::vector<int32_t> vert_ptr;    // [Nv+1]
    std::vector<int32_t> vert_inc;    // e * 4 + v
};

Pattern build_pattern(int n_elements, int n_local, const int32_t *elem_nodes, int n_nodes,
                      int n_vertices, const int32_t *elem_vertices);

// Per block (i, j) the list of element-local blocks (e, a, b) with
// node_e(a) = i and node_e(b) = j, ordered by element.  Encoded as
// (e * n_local + a) * n_local + b.
void build_block_contributions(const Pattern &pat, int n_local, const int32_t *elem_nodes,
                               std::vector<int64_t> &ptr, std::vector<int32_t> &contrib,
                               std::vector<int32_t> &block_row);

} // namespace pfgpu
