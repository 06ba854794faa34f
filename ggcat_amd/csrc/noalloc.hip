// noalloc.hip -- the query batch entry point that allocates nothing.
//
// No counterpart in the reference.  It is for callers that cannot call
// hipMalloc (tight serving loops, code that must not allocate): the queries
// and every result live in caller-owned device memory, and the only device
// memory beyond them is `scratch`, a fixed array of this library holding the
// work cursor, the per-block newline counts and the query offsets.  Its
// address is looked up once per device when a context is created
// (noalloc_scratch_setup) and only that address is handed to kernels.
//
// Every call on a device shares that device's array: calls that overlap on
// one GPU (two contexts or two threads) race and must be serialised by the
// caller.  The array keeps the last call's contents; what must start at zero
// (the work cursor, the error word) is cleared by a kernel at every call.
//
// Included by api.hip (one translation unit).

namespace {

// Scratch layout (bytes):
//   [0, 256)                    header: u32 work cursor, u32 error word, u32 newline total
//   [256, 256 + 4 * NA_MAX_BLOCKS)  newline count of each 16 KiB block, scanned in place
//   [NA_OFF_AT, ...)            (n_queries + 1) u64 query start offsets
constexpr uint32_t NA_BLOCK_BYTES = 16384;  // bytes per counting block (256 threads x 64 bytes)
constexpr uint32_t NA_MAX_BLOCKS = 16384;  // => at most 256 MiB of query bytes per call
constexpr uint64_t NA_OFF_AT = 256 + 4ull * NA_MAX_BLOCKS;
constexpr uint32_t NA_MAX_QUERIES = 500000;
constexpr uint64_t NA_SCRATCH_BYTES = 4ull << 20;
static_assert(NA_OFF_AT + 8ull * (NA_MAX_QUERIES + 1) <= NA_SCRATCH_BYTES, "scratch layout exceeds the array");

}  // namespace

static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NA_MAX_DEVICES] = {};

// Looks the array's device address up on `device` (which must be current).
// Called by ggcat_gpu_create; repeated calls on a device are no-ops.
static int noalloc_scratch_setup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return GGCAT_GPU_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return GGCAT_GPU_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return GGCAT_GPU_EHIP;
    }
    g_scratch_ptr[device] = p;
    return GGCAT_GPU_OK;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

__global__ void na_clear_kernel(uint32_t* header) {
    if (threadIdx.x < 64) header[threadIdx.x] = 0;
}

// Block b counts the '\n' bytes of [b * NA_BLOCK_BYTES, ...) (coalesced byte loads).
__global__ __launch_bounds__(256) void na_count_kernel(const uint8_t* seq, uint64_t n, uint32_t* block_nl) {
    __shared__ uint32_t s_part[4];
    const uint64_t base = (uint64_t)blockIdx.x * NA_BLOCK_BYTES;
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < NA_BLOCK_BYTES; i += 256) {
        const uint64_t p = base + i;
        if (p < n && seq[p] == '\n') c++;
    }
    for (uint32_t d = 32; d; d >>= 1) c += __shfl_down(c, d);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) block_nl[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

// One workgroup of 1024 threads: exclusive scan of nb <= NA_MAX_BLOCKS block
// counts in place; header[2] = total.
__global__ __launch_bounds__(1024) void na_scan_kernel(uint32_t* block_nl, uint32_t nb, uint32_t* header) {
    __shared__ uint32_t s_wave[16];
    constexpr uint32_t PER = NA_MAX_BLOCKS / 1024;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t v[PER], mine = 0;
    for (uint32_t q = 0; q < PER; q++) {
        const uint32_t i = t * PER + q;
        v[q] = i < nb ? block_nl[i] : 0u;
        mine += v[q];
    }
    uint32_t inc = mine;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t q = 0; q < 16; q++) {
        if (q < wave) before += s_wave[q];
        all += s_wave[q];
    }
    uint32_t run = before + inc - mine;
    for (uint32_t q = 0; q < PER; q++) {
        const uint32_t i = t * PER + q;
        if (i < nb) block_nl[i] = run;
        run += v[q];
    }
    if (t == 0) header[2] = all;
}

// Block b writes off[r + 1] = position after the r-th '\n' for the newlines of
// its range (rank r from the scanned block base plus a block scan); off[0] = 0.
// A rank at or beyond n_queries sets header[1].
__global__ __launch_bounds__(256) void na_fill_kernel(const uint8_t* seq, uint64_t n, const uint32_t* block_base,
                                                      uint64_t* off, uint32_t n_queries, uint32_t* header) {
    __shared__ uint32_t s_wave[4];
    constexpr uint32_t PER = NA_BLOCK_BYTES / 256;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint64_t p0 = (uint64_t)blockIdx.x * NA_BLOCK_BYTES + (uint64_t)t * PER;
    uint32_t mine = 0;
    for (uint32_t i = 0; i < PER; i++)
        if (p0 + i < n && seq[p0 + i] == '\n') mine++;
    uint32_t inc = mine;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t before = block_base[blockIdx.x];
    for (uint32_t q = 0; q < wave; q++) before += s_wave[q];
    uint32_t r = before + inc - mine;
    for (uint32_t i = 0; i < PER && r < before + inc; i++) {
        if (p0 + i < n && seq[p0 + i] == '\n') {
            if (r < n_queries)
                off[r + 1] = p0 + i + 1;
            else
                atomicOr(&header[1], 1u);
            r++;
        }
    }
    if (blockIdx.x == 0 && t == 0) off[0] = 0;
}

// header[1] |= 2 when the newline total differs from n_queries or the input
// does not end with '\n'.
__global__ void na_check_kernel(const uint8_t* seq, uint64_t n, uint32_t n_queries, uint32_t* header) {
    if (threadIdx.x != 0) return;
    if (header[2] != n_queries || (n && seq[n - 1] != '\n')) atomicOr(&header[1], 2u);
}

// query_kmers[q] = length - (k - 1), saturating (crates/querier/src/lib.rs:138-150).
__global__ void na_qkmers_kernel(const uint64_t* off, uint32_t n_queries, uint32_t k, uint64_t* out) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_queries) return;
    const uint64_t len = off[q + 1] - off[q] - 1;
    out[q] = len >= k ? len - (k - 1) : 0;
}

}  // namespace

extern "C" int ggcat_gpu_query_device_noalloc(ggcat_gpu_ctx* ctx, const uint8_t* d_queries, uint64_t n_bytes,
                                              uint32_t n_queries, uint64_t* d_query_kmers, uint64_t* d_matched,
                                              uint32_t* d_color_counts) {
    if (!ctx || (!d_queries && n_bytes)) return GGCAT_GPU_EINVAL;
    if (!ctx->have_index) return fail(ctx, GGCAT_GPU_ESTATE, "no query index (call ggcat_gpu_index_build)");
    const uint32_t nc = ctx->colored ? ctx->n_color_ids : 0;
    if (n_queries && (!d_matched || (nc && !d_color_counts)))
        return fail(ctx, GGCAT_GPU_EINVAL, "null device output");
    if (n_queries > NA_MAX_QUERIES)
        return fail(ctx, GGCAT_GPU_EINVAL, "no-allocation query: more than 500000 queries in one call");
    if (n_bytes > (uint64_t)NA_MAX_BLOCKS * NA_BLOCK_BYTES)
        return fail(ctx, GGCAT_GPU_EINVAL, "no-allocation query: more than 256 MiB of query bytes in one call");
    if (n_bytes == 0 && n_queries) return fail(ctx, GGCAT_GPU_EINVAL, "queries must be n_queries records, each ending with '\\n'");
    HIPCHK(hipSetDevice(ctx->device));
    uint8_t* S = static_cast<uint8_t*>(noalloc_scratch(ctx->device));
    if (!S) return fail(ctx, GGCAT_GPU_ESTATE, "scratch address of this device was not set up");
    uint32_t* header = reinterpret_cast<uint32_t*>(S);
    uint32_t* block_nl = reinterpret_cast<uint32_t*>(S + 256);
    uint64_t* off = reinterpret_cast<uint64_t*>(S + NA_OFF_AT);
    hipLaunchKernelGGL(na_clear_kernel, dim3(1), dim3(64), 0, ctx->stream, header);
    HIPCHK(hipGetLastError());
    if (n_queries == 0 && n_bytes == 0) return GGCAT_GPU_OK;
    const uint32_t nb = (uint32_t)((n_bytes + NA_BLOCK_BYTES - 1) / NA_BLOCK_BYTES);
    hipLaunchKernelGGL(na_count_kernel, dim3(nb), dim3(256), 0, ctx->stream, d_queries, n_bytes, block_nl);
    hipLaunchKernelGGL(na_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, block_nl, nb, header);
    hipLaunchKernelGGL(na_fill_kernel, dim3(nb), dim3(256), 0, ctx->stream, d_queries, n_bytes, block_nl, off,
                       n_queries, header);
    hipLaunchKernelGGL(na_check_kernel, dim3(1), dim3(64), 0, ctx->stream, d_queries, n_bytes, n_queries, header);
    HIPCHK(hipGetLastError());
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, header + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (err) return fail(ctx, GGCAT_GPU_EINVAL, "queries must be n_queries records, each ending with '\\n'");
    if (d_query_kmers)
        hipLaunchKernelGGL(na_qkmers_kernel, dim3((n_queries + 255) / 256), dim3(256), 0, ctx->stream, off, n_queries,
                           ctx->P.k, d_query_kmers);
    QueryArgs Q = query_args(ctx);
    Q.seq = d_queries;
    Q.qoff = off;
    Q.n_queries = n_queries;
    Q.matched = d_matched;
    Q.color_counts = d_color_counts;
    Q.work = header;  // header[0]: work cursor, zeroed above
    const uint32_t grid = std::min<uint32_t>(n_queries, ctx->n_cu * 4);
    hipLaunchKernelGGL(query_kernel, dim3(grid), dim3(256), (size_t)nc * 4 + 16, ctx->stream, Q);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GGCAT_GPU_OK;
}
