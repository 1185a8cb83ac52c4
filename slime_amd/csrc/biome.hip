// Biome filter of slime-cluster candidates and biome point queries on the
// GPU, behind the reference's mcgpu_modern_ctx_* C ABI
// (MinecraftGPUNoise.cu:978-1105, 1837-1986).
//
// Filter: a candidate is a center chunk; its disk's slime chunks (221-cell
// disk, Java-exact test) each hold 4 x 4 quart columns; flags bit 0 asks for
// Deep Dark (biome 183) at quart y -16..0 in any of them, bit 1 for Mushroom
// Fields (14) at quart y 16. One workgroup takes candidates in turn:
//   1. its threads test the 17 x 17 chunks and compact the slime chunks in LDS;
//   2. each thread takes quart columns: the five climate noises (52 double
//      Perlin octave samples) once per column, then per requested y layer the
//      depth and the nearest-leaf search of the parameter tree;
//   3. a found biome sets the candidate's mask bit in LDS; columns stop once
//      every requested bit is set.
// The climate state (46 octaves, 14.9 KB) is built on the host from the world
// seed and copied into each workgroup's LDS once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/mcgpu.h"
#include "biome.h"
#include "slime_math.h"

namespace slime {
namespace biome {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ void load_climate(Climate& dst, const Climate* src) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(&dst);
    for (int i = threadIdx.x; i < (int)(sizeof(Climate) / 4); i += blockDim.x) d[i] = s[i];
}

static_assert(sizeof(Climate) % 4 == 0, "climate copied as words");

// LDS of a filtering workgroup besides its climate.
struct FilterLds {
    int32_t params[2 * kTreeParamCount];  // the tree's intervals (read with divergent indices)
    uint16_t chunks[kDiskCells];  // slime chunks of the disk (index 17 * (dx+8) + dz+8)
    int nchunks;
    unsigned int mask;
};

// The leaves of the filter's two biomes (host: found in the tree once).
struct TargetLeaves {
    int32_t deep[4], ndeep;
    int32_t mush[4], nmush;
};

TargetLeaves find_targets() {
    TargetLeaves t{};
    for (int i = 0; i < kTreeLen; ++i) {
        if ((kTreeNodes[i] >> 56) != 0xFFu) continue;
        const int b = (int)((kTreeNodes[i] >> 48) & 0xFFu);
        if (b == kDeepDark && t.ndeep < 4) t.deep[t.ndeep++] = i;
        if (b == kMushroomFields && t.nmush < 4) t.mush[t.nmush++] = i;
    }
    return t;
}

// Whether the search at np can return one of `leaves`: not if every one of
// them lies strictly farther than the known leaf `hint` (the nearest leaf is
// no farther than hint). A biome with no leaf in the tree is never found.
__device__ __forceinline__ bool may_be(const int32_t np[6], const uint64_t* nodes,
                                       const int32_t* params, int hint, const int32_t* leaves,
                                       int nleaves) {
    if (nleaves == 0) return false;
    if (hint < 0) return true;
    const uint64_t known = node_distance(np, nodes, params, hint);
    for (int i = 0; i < nleaves; ++i)
        if (node_distance(np, nodes, params, leaves[i]) <= known) return true;
    return false;
}

// Whether layer yi of a column is the biome of `leaves`, searching only when
// a target leaf could be the nearest; hint carries the last leaf found.
__device__ __forceinline__ bool layer_is(const Column& col, int32_t yi, const uint64_t* nodes,
                                        const int32_t* params, int32_t len, int& hint,
                                        const int32_t* leaves, int nleaves, int biome) {
    const int32_t np[6] = {col.t, col.h, col.c, col.e, depth_at(col.offset, yi), col.w};
    if (!may_be(np, nodes, params, hint, leaves, nleaves)) return false;
    hint = tree_leaf(np, nodes, params, len, hint);
    return leaf_biome(nodes, hint) == biome;
}

__device__ __forceinline__ void load_params(int32_t* dst) {
    for (int i = threadIdx.x; i < 2 * kTreeParamCount; i += blockDim.x)
        dst[i] = kTreeParams[i >> 1][i & 1];
}

// The mask of one candidate, by the whole workgroup (every thread calls it;
// the caller synchronises before the first call when the LDS was in use).
__device__ __forceinline__ unsigned int filter_candidate(const Climate& st, FilterLds& f,
                                                         const uint64_t* nodes,
                                                         const int32_t* params, int32_t len,
                                                         const TargetLeaves& tl,
                                                         int64_t world_seed, int32_t cx,
                                                         int32_t cz, int32_t flags, int& hint) {
    if (threadIdx.x == 0) {
        f.nchunks = 0;
        f.mask = 0u;
    }
    __syncthreads();
    for (int ci = threadIdx.x; ci < 17 * 17; ci += blockDim.x) {
        const int dx = ci / 17 - 8, dz = ci % 17 - 8;
        if (dx * dx + dz * dz <= 68 &&
            is_slime(world_seed, (int32_t)((uint32_t)cx + (uint32_t)dx),
                     (int32_t)((uint32_t)cz + (uint32_t)dz))) {
            const int slot = atomicAdd(&f.nchunks, 1);
            f.chunks[slot] = (uint16_t)ci;
        }
    }
    __syncthreads();
    const int columns = f.nchunks * 16;
    const unsigned int want = (unsigned int)flags;
    for (int item = threadIdx.x; item < columns; item += blockDim.x) {
        if ((__atomic_load_n(&f.mask, __ATOMIC_RELAXED) & want) == want) break;
        const int ci = f.chunks[item >> 4], q = item & 15;
        const int dx = ci / 17 - 8, dz = ci % 17 - 8;
        const int32_t qx = (int32_t)(((uint32_t)cx + (uint32_t)dx) * 4u + (uint32_t)(q >> 2));
        const int32_t qz = (int32_t)(((uint32_t)cz + (uint32_t)dz) * 4u + (uint32_t)(q & 3));
        const Column col = column_at(st, qx, qz);
        unsigned int found = 0;
        // hint: the thread's last leaf, a bound for every later layer and column
        if ((want & 2u) && layer_is(col, 16, nodes, params, len, hint, tl.mush, tl.nmush,
                                    kMushroomFields))
            found |= 2u;
        if (want & 1u) {
            for (int32_t y = 0; y >= -16; --y) {
                if (layer_is(col, y, nodes, params, len, hint, tl.deep, tl.ndeep, kDeepDark)) {
                    found |= 1u;
                    break;
                }
            }
        }
        if (found) atomicOr(&f.mask, found);
    }
    __syncthreads();
    return f.mask;
}

__global__ __launch_bounds__(kThreads) void filter_kernel(const Climate* climate,
                                                          const uint64_t* nodes, int32_t len,
                                                          TargetLeaves tl, int64_t world_seed,
                                                          const int32_t* ccx, const int32_t* ccz,
                                                          int32_t count, int32_t flags,
                                                          int32_t* mask_out) {
    __shared__ Climate st;
    __shared__ FilterLds f;
    load_climate(st, climate);
    load_params(f.params);
    int hint = -1;
    for (int cand = blockIdx.x; cand < count; cand += gridDim.x) {
        __syncthreads();  // the previous candidate's LDS is no longer read
        const unsigned int m = filter_candidate(st, f, nodes, f.params, len, tl, world_seed,
                                                ccx[cand], ccz[cand], flags, hint);
        if (threadIdx.x == 0) mask_out[cand] = (int32_t)m;
    }
}

// Candidates of many worlds: world w's candidates are [first[w], first[w+1]),
// its climate climates[w]. A workgroup takes a contiguous run of candidates,
// so it reloads its LDS climate only where the run crosses into a new world.
__global__ __launch_bounds__(kThreads) void filter_multi_kernel(
    const Climate* climates, const int64_t* seeds, const int32_t* first, int32_t nworlds,
    const uint64_t* nodes, int32_t len, TargetLeaves tl, const int32_t* ccx, const int32_t* ccz,
    int32_t count, int32_t per_block, int32_t flags, int32_t* mask_out) {
    __shared__ Climate st;
    __shared__ FilterLds f;
    load_params(f.params);
    const int begin = blockIdx.x * per_block;
    const int end = min(count, begin + per_block);
    int world = -1;
    int hint = -1;
    for (int cand = begin; cand < end; ++cand) {
        int w = world < 0 ? 0 : world;
        while (w + 1 < nworlds && first[w + 1] <= cand) ++w;
        __syncthreads();  // the previous candidate's LDS is no longer read
        if (w != world) {
            load_climate(st, climates + w);
            world = w;
        }
        const unsigned int m = filter_candidate(st, f, nodes, f.params, len, tl, seeds[w],
                                                ccx[cand], ccz[cand], flags, hint);
        if (threadIdx.x == 0) mask_out[cand] = (int32_t)m;
    }
}

__global__ __launch_bounds__(kThreads) void query_kernel(const Climate* climate,
                                                         const uint64_t* nodes, int32_t len,
                                                         const int32_t* xs, const int32_t* ys,
                                                         const int32_t* zs, int32_t count,
                                                         int32_t* out) {
    __shared__ Climate st;
    __shared__ int32_t params[2 * kTreeParamCount];
    load_climate(st, climate);
    load_params(params);
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const Column col = column_at(st, xs[i], zs[i]);
        int hint = -1;
        out[i] = biome_at_column(col, ys[i], nodes, params, len, hint);
    }
}

// Per-column climate (six int32: t, h, c, e, w, offset bits): exposed for the
// parity tests of the noise stage alone.
__global__ __launch_bounds__(kThreads) void column_kernel(const Climate* climate,
                                                          const int32_t* xs, const int32_t* zs,
                                                          int32_t count, int32_t* out6) {
    __shared__ Climate st;
    load_climate(st, climate);
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const Column col = column_at(st, xs[i], zs[i]);
        int32_t* o = out6 + (size_t)i * 6;
        o[0] = col.t;
        o[1] = col.h;
        o[2] = col.c;
        o[3] = col.e;
        o[4] = col.w;
        union {
            float f;
            int32_t i;
        } u;
        u.f = col.offset;
        o[5] = u.i;
    }
}

}  // namespace

// A world's biome source on one device: climate state, tree and I/O buffers.
struct Context {
    int device = -1;
    int32_t capacity = 0;
    int64_t world_seed = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Climate* d_climate = nullptr;
    uint64_t* d_nodes = nullptr;
    int32_t *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_out = nullptr;
    float last_ms = 0.f;
    TargetLeaves targets = find_targets();

    ~Context() {
        if (device >= 0) (void)hipSetDevice(device);
        for (void* p : {(void*)d_climate, (void*)d_nodes, (void*)d_x, (void*)d_y, (void*)d_z,
                        (void*)d_out})
            if (p) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    bool init(uint64_t seed, bool large, int32_t cap) {
        if (hipGetDevice(&device) != hipSuccess) return false;
        capacity = cap;
        world_seed = (int64_t)seed;
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) ==
                hipSuccess &&
            n > 0)
            cus = n;
        Climate* host = new (std::nothrow) Climate;
        if (!host) return false;
        std::memset(host, 0, sizeof(Climate));
        climate_init(*host, seed, large);
        const size_t ni = sizeof(int32_t) * (size_t)cap;
        const bool ok =
            hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess &&
            hipMalloc(&d_climate, sizeof(Climate)) == hipSuccess &&
            hipMalloc(&d_nodes, sizeof(kTreeNodes)) == hipSuccess &&
            hipMalloc(&d_x, ni) == hipSuccess && hipMalloc(&d_y, ni) == hipSuccess &&
            hipMalloc(&d_z, ni) == hipSuccess && hipMalloc(&d_out, 6 * ni) == hipSuccess &&
            hipMemcpy(d_climate, host, sizeof(Climate), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(d_nodes, kTreeNodes, sizeof(kTreeNodes), hipMemcpyHostToDevice) ==
                hipSuccess;
        delete host;
        return ok;
    }

    int filter(const int32_t* cx, const int32_t* cz, int32_t count, int32_t flags, int32_t* out) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        const int grid = (int)std::min<int64_t>(count, (int64_t)cus * 4);
        if (hipMemcpyAsync(d_x, cx, ni, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(d_z, cz, ni, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipEventRecord(ev0, stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        hipLaunchKernelGGL(filter_kernel, dim3(grid), dim3(kThreads), 0, stream, d_climate,
                           d_nodes, kTreeLen, targets, world_seed, d_x, d_z, count, flags, d_out);
        if (hipGetLastError() != hipSuccess || hipEventRecord(ev1, stream) != hipSuccess ||
            hipMemcpyAsync(out, d_out, ni, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess ||
            hipEventElapsedTime(&last_ms, ev0, ev1) != hipSuccess)
            return MCGPU_ERR_CUDA;
        return MCGPU_OK;
    }

    int query(const int32_t* xs, const int32_t* ys, const int32_t* zs, int32_t count,
              int32_t* out) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        const int grid = (int)std::min<int64_t>((count + kThreads - 1) / kThreads, (int64_t)cus * 4);
        if (hipMemcpyAsync(d_x, xs, ni, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(d_y, ys, ni, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(d_z, zs, ni, hipMemcpyHostToDevice, stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        hipLaunchKernelGGL(query_kernel, dim3(grid), dim3(kThreads), 0, stream, d_climate, d_nodes,
                           kTreeLen, d_x, d_y, d_z, count, d_out);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, d_out, ni, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        return MCGPU_OK;
    }

    int columns(const int32_t* xs, const int32_t* zs, int32_t count, int32_t* out6) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        const int grid = (int)std::min<int64_t>((count + kThreads - 1) / kThreads, (int64_t)cus * 4);
        if (hipMemcpyAsync(d_x, xs, ni, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(d_z, zs, ni, hipMemcpyHostToDevice, stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        hipLaunchKernelGGL(column_kernel, dim3(grid), dim3(kThreads), 0, stream, d_climate, d_x,
                           d_z, count, d_out);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out6, d_out, 6 * ni, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        return MCGPU_OK;
    }
};

// Filter of candidates of many worlds in one launch (see filter_multi_kernel).
int filter_multi(int32_t nworlds, const uint64_t* seeds, const int32_t* first, const int32_t* cx,
                 const int32_t* cz, int32_t flags, bool large, int32_t* out, float* ms) {
    const int32_t count = first[nworlds];
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return MCGPU_ERR_CUDA;
    std::vector<Climate> host((size_t)nworlds);
    std::vector<int64_t> sseed((size_t)nworlds);
    for (int32_t w = 0; w < nworlds; ++w) {
        std::memset(&host[w], 0, sizeof(Climate));
        climate_init(host[w], seeds[w], large);
        sseed[w] = (int64_t)seeds[w];
    }
    Climate* d_clim = nullptr;
    int64_t* d_seeds = nullptr;
    int32_t *d_first = nullptr, *d_x = nullptr, *d_z = nullptr, *d_out = nullptr;
    uint64_t* d_nodes = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t ni = sizeof(int32_t) * (size_t)count;
    // about 4 workgroups per CU's worth of runs, each a contiguous range
    const int32_t blocks_wanted = std::max(1, cus * 4);
    const int32_t per_block = std::max<int32_t>(1, (count + blocks_wanted - 1) / blocks_wanted);
    const int32_t grid = (count + per_block - 1) / per_block;
    bool ok = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              hipMalloc(&d_clim, sizeof(Climate) * (size_t)nworlds) == hipSuccess &&
              hipMalloc(&d_seeds, sizeof(int64_t) * (size_t)nworlds) == hipSuccess &&
              hipMalloc(&d_first, sizeof(int32_t) * (size_t)(nworlds + 1)) == hipSuccess &&
              hipMalloc(&d_nodes, sizeof(kTreeNodes)) == hipSuccess &&
              hipMalloc(&d_x, ni) == hipSuccess && hipMalloc(&d_z, ni) == hipSuccess &&
              hipMalloc(&d_out, ni) == hipSuccess &&
              hipMemcpyAsync(d_clim, host.data(), sizeof(Climate) * (size_t)nworlds,
                             hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_seeds, sseed.data(), sizeof(int64_t) * (size_t)nworlds,
                             hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_first, first, sizeof(int32_t) * (size_t)(nworlds + 1),
                             hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_nodes, kTreeNodes, sizeof(kTreeNodes), hipMemcpyHostToDevice,
                             stream) == hipSuccess &&
              hipMemcpyAsync(d_x, cx, ni, hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_z, cz, ni, hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipEventRecord(e0, stream) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(filter_multi_kernel, dim3(grid), dim3(kThreads), 0, stream, d_clim,
                           d_seeds, d_first, nworlds, d_nodes, kTreeLen, find_targets(), d_x, d_z,
                           count, per_block, flags, d_out);
        ok = hipGetLastError() == hipSuccess && hipEventRecord(e1, stream) == hipSuccess &&
             hipMemcpyAsync(out, d_out, ni, hipMemcpyDeviceToHost, stream) == hipSuccess &&
             hipStreamSynchronize(stream) == hipSuccess &&
             hipEventElapsedTime(ms, e0, e1) == hipSuccess;
    }
    for (void* p : {(void*)d_clim, (void*)d_seeds, (void*)d_first, (void*)d_nodes, (void*)d_x,
                    (void*)d_z, (void*)d_out})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (stream) (void)hipStreamDestroy(stream);
    return ok ? MCGPU_OK : MCGPU_ERR_CUDA;
}

}  // namespace biome
}  // namespace slime

using slime::biome::Context;

namespace {
float g_multi_ms = 0.f;
}

extern "C" {

int mcgpu_api_version(void) { return 1; }

const char* mcgpu_build_name(void) { return "slime_amd biome filter (gfx950, HIP)"; }

void* mcgpu_modern_ctx_create_ex(int32_t version_code, uint64_t world_seed, int32_t flags,
                                 int32_t capacity) {
    // the parameter tree here is the 26.2 one: older versions are refused
    if (capacity <= 0 || version_code < 2602 || (flags & ~1) != 0) return nullptr;
    Context* ctx = new (std::nothrow) Context;
    if (!ctx) return nullptr;
    if (!ctx->init(world_seed, (flags & 1) != 0, capacity)) {
        delete ctx;
        (void)hipGetLastError();
        return nullptr;
    }
    return ctx;
}

void* mcgpu_modern_ctx_create(int32_t version_code, uint64_t world_seed, int32_t capacity) {
    return mcgpu_modern_ctx_create_ex(version_code, world_seed, 0, capacity);
}

void mcgpu_modern_ctx_destroy(void* handle) { delete static_cast<Context*>(handle); }

int mcgpu_modern_ctx_filter_candidates(void* handle, const int32_t* center_chunk_x,
                                       const int32_t* center_chunk_z, int32_t count,
                                       int32_t flags, int32_t* mask_out) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || !center_chunk_x || !center_chunk_z || !mask_out || count <= 0 ||
        count > ctx->capacity || (flags & ~3) != 0 || flags == 0)
        return MCGPU_ERR_ARG;
    const int rc = ctx->filter(center_chunk_x, center_chunk_z, count, flags, mask_out);
    if (rc != MCGPU_OK) (void)hipGetLastError();
    return rc;
}

int mcgpu_modern_ctx_query(void* handle, const int32_t* xs, const int32_t* ys, const int32_t* zs,
                           int32_t count, int32_t* biome_out) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || !xs || !ys || !zs || !biome_out || count <= 0 || count > ctx->capacity)
        return MCGPU_ERR_ARG;
    const int rc = ctx->query(xs, ys, zs, count, biome_out);
    if (rc != MCGPU_OK) (void)hipGetLastError();
    return rc;
}

int32_t mcgpu_modern_recommended_filter_batch(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
        (void)hipGetLastError();
        return 128;
    }
    return cus * 16;  // candidates per call that keep every CU's workgroups busy
}

int mcgpu_modern_ctx_columns(void* handle, const int32_t* xs, const int32_t* zs, int32_t count,
                             int32_t* out6) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || !xs || !zs || !out6 || count <= 0 || count > ctx->capacity) return MCGPU_ERR_ARG;
    const int rc = ctx->columns(xs, zs, count, out6);
    if (rc != MCGPU_OK) (void)hipGetLastError();
    return rc;
}

int mcgpu_filter_multi(int32_t nworlds, const uint64_t* world_seeds, const int32_t* first,
                       const int32_t* center_chunk_x, const int32_t* center_chunk_z,
                       int32_t flags, int32_t large_biomes, int32_t* mask_out) {
    if (nworlds <= 0 || !world_seeds || !first || !center_chunk_x || !center_chunk_z ||
        !mask_out || (flags & ~3) != 0 || flags == 0 || first[0] != 0)
        return MCGPU_ERR_ARG;
    for (int32_t w = 0; w < nworlds; ++w)
        if (first[w + 1] < first[w]) return MCGPU_ERR_ARG;
    if (first[nworlds] == 0) return MCGPU_OK;
    const int rc = slime::biome::filter_multi(nworlds, world_seeds, first, center_chunk_x,
                                              center_chunk_z, flags, large_biomes != 0, mask_out,
                                              &g_multi_ms);
    if (rc != MCGPU_OK) (void)hipGetLastError();
    return rc;
}

float mcgpu_filter_multi_last_ms(void) { return g_multi_ms; }

float mcgpu_modern_ctx_last_filter_ms(void* handle) {
    Context* ctx = static_cast<Context*>(handle);
    return ctx ? ctx->last_ms : -1.f;
}

}  // extern "C"
