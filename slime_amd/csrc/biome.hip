// Biome filter of slime-cluster candidates and biome / climate point queries
// on the GPU, behind the reference's mcgpu_* C ABI (MinecraftGPUNoise.cu:
// 424-1105 kernels and primitives, 1405-2018 contexts and queries).
//
// Filter: a candidate is a center chunk; its disk's slime chunks (221-cell
// disk, Java-exact test) each hold 4 x 4 quart columns; flags bit 0 asks for
// Deep Dark (biome 183) at quart y -16..0 in any of them, bit 1 for Mushroom
// Fields (14) at quart y 16. One workgroup takes candidates in turn:
//   1. its threads test the 17 x 17 chunks and compact the slime chunks in LDS;
//   2. each thread takes quart columns: the five climate noises (52 double
//      Perlin octave samples) once per column, then per requested y layer the
//      depth and the nearest-leaf search of the version's parameter tree;
//   3. a found biome sets the candidate's mask bit in LDS; columns stop once
//      every requested bit is set.
// Besides the reference's heights, the filter takes a range of block heights
// (the wider sweep y -64..320 step 4, say) at which both biomes are looked
// for, one tree search per height, and optionally every chunk of the disk
// (flags bit 2) instead of its slime chunks.
// The climate state (46 octaves, 14.9 KB) is built on the host from the world
// seed and copied into each workgroup's LDS once; so are the tree's intervals.
// Kernels that search a tree are instantiated for both node layouts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
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
static_assert(sizeof(Perlin) % 4 == 0, "Perlin state copied as words");

// A parameter tree on the device.
struct TreeArgs {
    const uint64_t* nodes;
    const int32_t* params;  // (lo, hi) pairs
    int32_t len;
    int32_t nparams;
};

__device__ __forceinline__ void load_params(int32_t* dst, const TreeArgs& t) {
    for (int i = threadIdx.x; i < 2 * t.nparams; i += blockDim.x) dst[i] = t.params[i];
}

// LDS of a filtering workgroup besides its climate.
struct FilterLds {
    int32_t params[2 * kTreeMaxParams];  // the tree's intervals (read with divergent indices)
    uint16_t chunks[kDiskCells];  // slime chunks of the disk (index 17 * (dx+8) + dz+8)
    int nchunks;
    unsigned int mask;
};

// The leaves of the filter's two biomes (host: found in the tree once).
constexpr int kMaxTargetLeaves = 4;
struct TargetLeaves {
    int32_t deep[kMaxTargetLeaves], ndeep;
    int32_t mush[kMaxTargetLeaves], nmush;
};

// Returns false if a biome has more leaves than the list holds.
bool find_targets(const BiomeTreeDef& tree, TargetLeaves& t) {
    t = TargetLeaves{};
    for (int i = 0; i < tree.len; ++i) {
        if ((tree.nodes[i] >> 56) != 0xFFu) continue;
        const int b = (int)((tree.nodes[i] >> 48) & 0xFFu);
        if (b == kDeepDark) {
            if (t.ndeep == kMaxTargetLeaves) return false;
            t.deep[t.ndeep++] = i;
        }
        if (b == kMushroomFields) {
            if (t.nmush == kMaxTargetLeaves) return false;
            t.mush[t.nmush++] = i;
        }
    }
    return true;
}

// Stops a search at the first kept leaf that is not one of the targets (the
// leaves held by value: constant indices keep them in registers; -1 marks an
// unused entry).
template <int N>
struct NotTarget {
    int32_t leaf[N];
    __device__ bool operator()(int l) const {
        bool hit = false;
#pragma unroll
        for (int i = 0; i < N; ++i) hit |= leaf[i] == l;
        return !hit;
    }
};

template <int N>
__device__ __forceinline__ NotTarget<N> targets_of(const int32_t* leaves, int n) {
    NotTarget<N> t;
#pragma unroll
    for (int i = 0; i < N; ++i) t.leaf[i] = i < n ? leaves[i] : -1;
    return t;
}

// The leaf the search from the root at layer yi of a column returns if it is
// one of the targets, else -1. Not a target if every one of them lies
// strictly farther than `hint`, a leaf found before (the nearest leaf is no
// farther). Otherwise the search runs with the bound (nearest target's
// distance + 1) and stops at the first kept leaf that is not a target: that
// leaf is no farther than every target and comes first in visiting order
// among the leaves that near, so the result cannot be a target. If the search
// ends without stopping, every leaf it kept was a target and the last is the
// result. `hint` gets the last leaf kept (a bound for the next layers).
template <int F, int N>
__device__ __forceinline__ int layer_target(const Column& col, int32_t yi, const uint64_t* nodes,
                                            const int32_t* params, int32_t len, int& hint,
                                            const NotTarget<N>& other) {
    const int32_t np[6] = {col.t, col.h, col.c, col.e, depth_at(col.offset, yi), col.w};
    uint64_t target = ~0ull;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (other.leaf[i] >= 0) {
            const uint64_t d = node_distance(np, nodes, params, other.leaf[i]);
            target = d < target ? d : target;
        }
    }
    if (target == ~0ull) return -1;
    if (hint >= 0 && target > node_distance(np, nodes, params, hint)) return -1;
    bool stopped = false;
    const int leaf =
        tree_nearest<F, 0>(np, nodes, params, len, 0, -1, target + 1, other, stopped);
    if (leaf >= 0) hint = leaf;
    return stopped ? -1 : leaf;
}

// The heights a filter samples. n == 0: the reference's set (Deep Dark at
// quart y -16..0, Mushroom Fields at quart y 16). Otherwise every requested
// biome at block heights y0, y0 + step, ... (n of them), sampled at quart y
// floor(y / 4), one search per height for all biomes still wanted.
struct Levels {
    int32_t y0, step, n;
};

__device__ __forceinline__ int32_t floor_quart(int32_t y) { return y >> 2; }

// The mask of one candidate, by the whole workgroup (every thread calls it;
// the caller synchronises before the first call when the LDS was in use).
// L: the level-range form (lv.n > 0), a kernel of its own so that the
// reference's form keeps its register budget.
template <int F, bool L>
__device__ __forceinline__ unsigned int filter_candidate(const Climate& st, FilterLds& f,
                                                         const uint64_t* nodes, int32_t len,
                                                         const TargetLeaves& tl,
                                                         int64_t world_seed, int32_t cx,
                                                         int32_t cz, int32_t flags,
                                                         const Levels& lv, int& hint) {
    if (threadIdx.x == 0) {
        f.nchunks = 0;
        f.mask = 0u;
    }
    __syncthreads();
    const bool all_chunks = (flags & 4) != 0;
    for (int ci = threadIdx.x; ci < 17 * 17; ci += blockDim.x) {
        const int dx = ci / 17 - 8, dz = ci % 17 - 8;
        if (dx * dx + dz * dz <= 68 &&
            (all_chunks || is_slime(world_seed, (int32_t)((uint32_t)cx + (uint32_t)dx),
                                    (int32_t)((uint32_t)cz + (uint32_t)dz)))) {
            const int slot = atomicAdd(&f.nchunks, 1);
            f.chunks[slot] = (uint16_t)ci;
        }
    }
    __syncthreads();
    const int columns = f.nchunks * 16;
    const unsigned int want = (unsigned int)flags & 3u;
    const int32_t* params = f.params;
    const NotTarget<kMaxTargetLeaves> deep = targets_of<kMaxTargetLeaves>(tl.deep, tl.ndeep);
    const NotTarget<kMaxTargetLeaves> mush = targets_of<kMaxTargetLeaves>(tl.mush, tl.nmush);
    for (int item = threadIdx.x; item < columns; item += blockDim.x) {
        // bits another column already found need no more checks
        unsigned int need = want & ~__atomic_load_n(&f.mask, __ATOMIC_RELAXED);
        if (need == 0) break;
        const int ci = f.chunks[item >> 4], q = item & 15;
        const int dx = ci / 17 - 8, dz = ci % 17 - 8;
        const int32_t qx = (int32_t)(((uint32_t)cx + (uint32_t)dx) * 4u + (uint32_t)(q >> 2));
        const int32_t qz = (int32_t)(((uint32_t)cz + (uint32_t)dz) * 4u + (uint32_t)(q & 3));
        const Column col = column_at(st, qx, qz);
        unsigned int found = 0;
        // hint: the thread's last leaf, a bound for every later layer and column
        if constexpr (!L) {
            if ((need & 2u) && layer_target<F>(col, 16, nodes, params, len, hint, mush) >= 0)
                found |= 2u;
            if (need & 1u) {
                for (int32_t y = -16; y <= 0; ++y) {  // deepest first: Deep Dark lies deep
                    if (layer_target<F>(col, y, nodes, params, len, hint, deep) >= 0) {
                        found |= 1u;
                        break;
                    }
                }
            }
        } else {
            int32_t prev = 0;
            for (int32_t i = 0; i < lv.n && need != 0; ++i) {
                const int32_t yi = floor_quart(lv.y0 + i * lv.step);
                if (i > 0 && yi == prev) continue;  // steps below 4 blocks share quarts
                prev = yi;
                NotTarget<2 * kMaxTargetLeaves> both;
#pragma unroll
                for (int k = 0; k < kMaxTargetLeaves; ++k) {
                    both.leaf[k] = (need & 1u) ? deep.leaf[k] : -1;
                    both.leaf[kMaxTargetLeaves + k] = (need & 2u) ? mush.leaf[k] : -1;
                }
                const int l = layer_target<F>(col, yi, nodes, params, len, hint, both);
                if (l >= 0) {
                    const bool is_deep = !deep(l);  // deep(l) is true for other leaves
                    const unsigned int bit = is_deep ? 1u : 2u;
                    found |= bit;
                    need &= ~bit;
                }
            }
        }
        if (found) atomicOr(&f.mask, found);
    }
    __syncthreads();
    return f.mask;
}

template <int F, bool L>
__global__ __launch_bounds__(kThreads) void filter_kernel(const Climate* climate, TreeArgs tree,
                                                          TargetLeaves tl, int64_t world_seed,
                                                          const int32_t* ccx, const int32_t* ccz,
                                                          int32_t count, int32_t flags,
                                                          Levels lv, int32_t* mask_out) {
    __shared__ Climate st;
    __shared__ FilterLds f;
    load_climate(st, climate);
    load_params(f.params, tree);
    int hint = -1;
    for (int cand = blockIdx.x; cand < count; cand += gridDim.x) {
        __syncthreads();  // the previous candidate's LDS is no longer read
        const unsigned int m = filter_candidate<F, L>(st, f, tree.nodes, tree.len, tl, world_seed,
                                                   ccx[cand], ccz[cand], flags, lv, hint);
        if (threadIdx.x == 0) mask_out[cand] = (int32_t)m;
    }
}

// Candidates of many worlds: world w's candidates are [first[w], first[w+1]),
// its climate climates[w]. The candidates form a device work queue: a
// workgroup takes the next one with one atomic add on *next (zero at launch)
// until the list is exhausted, so that sites whose checks end early (a Deep
// Dark quart found in the first layers) do not leave their CUs idle while
// others still run; it reloads its LDS climate only when the world changes.
// Every workgroup leaves the loop once the counter passes count.
template <int F, bool L>
__global__ __launch_bounds__(kThreads) void filter_multi_kernel(
    const Climate* climates, const int64_t* seeds, const int32_t* first, int32_t nworlds,
    TreeArgs tree, TargetLeaves tl, const int32_t* ccx, const int32_t* ccz, int32_t count,
    unsigned int* next, int32_t flags, Levels lv, int32_t* mask_out) {
    __shared__ Climate st;
    __shared__ FilterLds f;
    __shared__ int s_cand;
    load_params(f.params, tree);
    int world = -1;
    int hint = -1;
    for (;;) {
        __syncthreads();  // the previous candidate's LDS is no longer read
        if (threadIdx.x == 0) s_cand = (int)atomicAdd(next, 1u);
        __syncthreads();
        const int cand = s_cand;
        if (cand >= count) break;
        // the candidate's world: the last w with first[w] <= cand
        int lo = 0, hi = nworlds - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= cand) lo = mid;
            else hi = mid - 1;
        }
        if (lo != world) {
            load_climate(st, climates + lo);
            world = lo;
        }
        const unsigned int m = filter_candidate<F, L>(st, f, tree.nodes, tree.len, tl,
                                                   seeds[lo], ccx[cand], ccz[cand], flags, lv,
                                                   hint);
        if (threadIdx.x == 0) mask_out[cand] = (int32_t)m;
    }
}

// Biome ids at quart points (MinecraftGPUNoise.cu:868-976).
template <int F>
__global__ __launch_bounds__(kThreads) void query_kernel(const Climate* climate, TreeArgs tree,
                                                         const int32_t* xs, const int32_t* ys,
                                                         const int32_t* zs, int32_t count,
                                                         int32_t* out) {
    __shared__ Climate st;
    __shared__ int32_t params[2 * kTreeMaxParams];
    load_climate(st, climate);
    load_params(params, tree);
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const Column col = column_at(st, xs[i], zs[i]);
        int hint = -1;
        out[i] = biome_at_column<F>(col, ys[i], tree.nodes, params, tree.len, hint);
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

// Climate at points, the reference's point-query layouts
// (MinecraftGPUNoise.cu:532-655, 811-866):
//   kRaw5    five doubles (t, h, c, e, w) at double coordinates;
//   kQuant5  the same quantized to int32;
//   kQuant6  six int32 (t, h, c, e, depth at quart y, w).
enum ClimateOut { kRaw5 = 0, kQuant5 = 1, kQuant6 = 2 };

template <typename Coord, int Mode>
__global__ __launch_bounds__(kThreads) void climate_kernel(const Climate* climate,
                                                           const Coord* xs, const int32_t* ys,
                                                           const Coord* zs, int32_t count,
                                                           void* out) {
    __shared__ Climate st;
    load_climate(st, climate);
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        double px, pz;
        shifted_point(st, (double)xs[i], (double)zs[i], px, pz);
        const double c = noise_sample(st, kContinentalness, px, 0.0, pz);
        const double e = noise_sample(st, kErosion, px, 0.0, pz);
        const double w = noise_sample(st, kWeirdness, px, 0.0, pz);
        const double t = noise_sample(st, kTemperature, px, 0.0, pz);
        const double h = noise_sample(st, kHumidity, px, 0.0, pz);
        if constexpr (Mode == kRaw5) {
            double* o = static_cast<double*>(out) + (size_t)i * 5;
            o[0] = t;
            o[1] = h;
            o[2] = c;
            o[3] = e;
            o[4] = w;
        } else if constexpr (Mode == kQuant5) {
            int32_t* o = static_cast<int32_t*>(out) + (size_t)i * 5;
            o[0] = quantize(t);
            o[1] = quantize(h);
            o[2] = quantize(c);
            o[3] = quantize(e);
            o[4] = quantize(w);
        } else {
            int32_t* o = static_cast<int32_t*>(out) + (size_t)i * 6;
            const float cf = (float)c, ef = (float)e, wf = (float)w;
            const float sv[4] = {cf, ef, peaks_and_valleys(wf), wf};
            o[0] = quantize(t);
            o[1] = quantize(h);
            o[2] = quantize(c);
            o[3] = quantize(e);
            o[4] = depth_at(spline_eval<0>(0, sv) + 0.015F, ys[i]);
            o[5] = quantize(w);
        }
    }
}

// The continentalness noise alone, unquantized, at (x, 0, z) without the
// shift (MinecraftGPUNoise.cu:491-517).
__global__ __launch_bounds__(kThreads) void continentalness_kernel(const Climate* climate,
                                                                   const double* xs,
                                                                   const double* zs,
                                                                   int32_t count, double* out) {
    __shared__ Climate st;
    load_climate(st, climate);
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        out[i] = noise_sample(st, kContinentalness, xs[i], 0.0, zs[i]);
}

// `steps` xoroshiro128++ outputs of each seed (MinecraftGPUNoise.cu:424-443).
__global__ __launch_bounds__(kThreads) void xoro_kernel(const uint64_t* seeds, int32_t count,
                                                        int32_t steps, uint64_t* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    Xoro s = xoro_from_seed(seeds[i]);
    uint64_t* o = out + (size_t)i * (size_t)steps;
    for (int32_t j = 0; j < steps; ++j) o[j] = xoro_next(s);
}

// One Perlin noise per query, seeded from its own seed
// (MinecraftGPUNoise.cu:445-461): the state is built in the thread.
__global__ __launch_bounds__(kThreads) void perlin_batch_kernel(const uint64_t* seeds,
                                                                const double* xs,
                                                                const double* ys,
                                                                const double* zs, int32_t count,
                                                                double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    Perlin p;
    perlin_init(p, xoro_from_seed(seeds[i]));
    out[i] = perlin_sample(p, xs[i], ys[i], zs[i]);
}

// One Perlin noise for every query (MinecraftGPUNoise.cu:463-489).
__global__ __launch_bounds__(kThreads) void perlin_fixed_kernel(const Perlin* state,
                                                                const double* xs,
                                                                const double* ys,
                                                                const double* zs, int32_t count,
                                                                double* out) {
    __shared__ Perlin p;
    const uint32_t* s = reinterpret_cast<const uint32_t*>(state);
    uint32_t* d = reinterpret_cast<uint32_t*>(&p);
    for (int k = threadIdx.x; k < (int)(sizeof(Perlin) / 4); k += blockDim.x) d[k] = s[k];
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        out[i] = perlin_sample(p, xs[i], ys[i], zs[i]);
}

int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0) {
        (void)hipGetLastError();
        return 256;
    }
    return n;
}

int grid_for(int32_t count, int cus) {
    return (int)std::max<int64_t>(
        1, std::min<int64_t>((count + kThreads - 1) / kThreads, (int64_t)cus * 4));
}

// Device copies of the trees, per device, made on first use and kept until
// the process ends (a few hundred KB).
struct DeviceTrees {
    int device = -1;
    uint64_t* nodes[kTreeCount] = {};
    int32_t* params[kTreeCount] = {};
};

std::mutex g_tree_mu;
std::vector<DeviceTrees> g_trees;

bool device_tree(int tree, TreeArgs& out) {
    int dev = 0;
    if (tree < 0 || tree >= kTreeCount || hipGetDevice(&dev) != hipSuccess) return false;
    std::lock_guard<std::mutex> lock(g_tree_mu);
    DeviceTrees* dt = nullptr;
    for (auto& d : g_trees)
        if (d.device == dev) dt = &d;
    if (!dt) {
        g_trees.emplace_back();
        dt = &g_trees.back();
        dt->device = dev;
    }
    const BiomeTreeDef& def = kTreeDefs[tree];
    if (!dt->nodes[tree]) {
        uint64_t* n = nullptr;
        int32_t* p = nullptr;
        const size_t nb = sizeof(uint64_t) * (size_t)def.len;
        const size_t pb = sizeof(int32_t) * 2 * (size_t)def.nparams;
        if (hipMalloc(&n, nb) != hipSuccess || hipMalloc(&p, pb) != hipSuccess ||
            hipMemcpy(n, def.nodes, nb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p, def.params, pb, hipMemcpyHostToDevice) != hipSuccess) {
            if (n) (void)hipFree(n);
            if (p) (void)hipFree(p);
            return false;
        }
        dt->nodes[tree] = n;
        dt->params[tree] = p;
    }
    out.nodes = dt->nodes[tree];
    out.params = dt->params[tree];
    out.len = def.len;
    out.nparams = def.nparams;
    return true;
}

}  // namespace

// A world's biome source on one device: climate state, parameter tree (or
// none for climate-only contexts) and I/O buffers of `capacity` points.
struct Context {
    int device = -1;
    int32_t capacity = 0;
    int64_t world_seed = 0;
    int cus = 256;
    int tree = -1;
    int family = 1;
    TreeArgs targs{};
    TargetLeaves targets{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Climate* d_climate = nullptr;
    // d_x / d_z hold an int32 or a double per point, d_out up to five
    // doubles (ten words) per point.
    int32_t *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_out = nullptr;
    float last_ms = 0.f;

    ~Context() {
        if (device >= 0) (void)hipSetDevice(device);
        for (void* p : {(void*)d_climate, (void*)d_x, (void*)d_y, (void*)d_z, (void*)d_out})
            if (p) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // MCGPU_OK, or MCGPU_ERR_ALLOC / MCGPU_ERR_CUDA.
    int init(uint64_t seed, bool large, int tree_index, int32_t cap) {
        if (hipGetDevice(&device) != hipSuccess) return MCGPU_ERR_CUDA;
        capacity = cap;
        world_seed = (int64_t)seed;
        cus = device_cus();
        tree = tree_index;
        if (tree >= 0) {
            family = kTreeDefs[tree].family;
            if (!find_targets(kTreeDefs[tree], targets) || !device_tree(tree, targs))
                return MCGPU_ERR_ALLOC;
        }
        Climate* host = new (std::nothrow) Climate;
        if (!host) return MCGPU_ERR_ALLOC;
        std::memset(host, 0, sizeof(Climate));
        climate_init(*host, seed, large);
        const size_t ni = sizeof(int32_t) * (size_t)cap;
        const bool alloc = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess &&
                           hipEventCreate(&ev0) == hipSuccess &&
                           hipEventCreate(&ev1) == hipSuccess &&
                           hipMalloc(&d_climate, sizeof(Climate)) == hipSuccess &&
                           hipMalloc(&d_x, 2 * ni) == hipSuccess &&
                           hipMalloc(&d_y, ni) == hipSuccess &&
                           hipMalloc(&d_z, 2 * ni) == hipSuccess &&
                           hipMalloc(&d_out, 10 * ni) == hipSuccess;
        const bool copied =
            alloc && hipMemcpy(d_climate, host, sizeof(Climate), hipMemcpyHostToDevice) ==
                         hipSuccess;
        delete host;
        return !alloc ? MCGPU_ERR_ALLOC : copied ? MCGPU_OK : MCGPU_ERR_CUDA;
    }

    int finish(void* out, size_t bytes) {
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        return MCGPU_OK;
    }

    bool upload(int32_t* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream) == hipSuccess;
    }

    int filter(const int32_t* cx, const int32_t* cz, int32_t count, int32_t flags,
               const Levels& lv, int32_t* out) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        const int grid = (int)std::min<int64_t>(count, (int64_t)cus * 4);
        if (!upload(d_x, cx, ni) || !upload(d_z, cz, ni) ||
            hipEventRecord(ev0, stream) != hipSuccess)
            return MCGPU_ERR_CUDA;
        auto* kernel = lv.n == 0 ? (family == 0 ? filter_kernel<0, false> : filter_kernel<1, false>)
                                 : (family == 0 ? filter_kernel<0, true> : filter_kernel<1, true>);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, stream, d_climate, targs,
                           targets, world_seed, d_x, d_z, count, flags, lv, d_out);
        if (hipGetLastError() != hipSuccess || hipEventRecord(ev1, stream) != hipSuccess ||
            finish(out, ni) != MCGPU_OK || hipEventElapsedTime(&last_ms, ev0, ev1) != hipSuccess)
            return MCGPU_ERR_CUDA;
        return MCGPU_OK;
    }

    int query(const int32_t* xs, const int32_t* ys, const int32_t* zs, int32_t count,
              int32_t* out) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        const int grid = grid_for(count, cus);
        if (!upload(d_x, xs, ni) || !upload(d_y, ys, ni) || !upload(d_z, zs, ni))
            return MCGPU_ERR_CUDA;
        if (family == 0)
            hipLaunchKernelGGL(query_kernel<0>, dim3(grid), dim3(kThreads), 0, stream, d_climate,
                               targs, d_x, d_y, d_z, count, d_out);
        else
            hipLaunchKernelGGL(query_kernel<1>, dim3(grid), dim3(kThreads), 0, stream, d_climate,
                               targs, d_x, d_y, d_z, count, d_out);
        return finish(out, ni);
    }

    int columns(const int32_t* xs, const int32_t* zs, int32_t count, int32_t* out6) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        if (!upload(d_x, xs, ni) || !upload(d_z, zs, ni)) return MCGPU_ERR_CUDA;
        hipLaunchKernelGGL(column_kernel, dim3(grid_for(count, cus)), dim3(kThreads), 0, stream,
                           d_climate, d_x, d_z, count, d_out);
        return finish(out6, 6 * ni);
    }

    // Climate at int32 points: five quantized values, or six with the depth
    // at quart y (ys != null).
    int climate_i32(const int32_t* xs, const int32_t* ys, const int32_t* zs, int32_t count,
                    int32_t* out) {
        const size_t ni = sizeof(int32_t) * (size_t)count;
        if (!upload(d_x, xs, ni) || !upload(d_z, zs, ni) || (ys && !upload(d_y, ys, ni)))
            return MCGPU_ERR_CUDA;
        if (ys)
            hipLaunchKernelGGL((climate_kernel<int32_t, kQuant6>), dim3(grid_for(count, cus)),
                               dim3(kThreads), 0, stream, d_climate, d_x, d_y, d_z, count,
                               (void*)d_out);
        else
            hipLaunchKernelGGL((climate_kernel<int32_t, kQuant5>), dim3(grid_for(count, cus)),
                               dim3(kThreads), 0, stream, d_climate, d_x, nullptr, d_z, count,
                               (void*)d_out);
        return finish(out, (ys ? 6 : 5) * ni);
    }

    // Climate at double points: five raw doubles, or five quantized int32.
    int climate_f64(const double* xs, const double* zs, int32_t count, bool quantized, void* out) {
        const size_t nd = sizeof(double) * (size_t)count;
        if (!upload(d_x, xs, nd) || !upload(d_z, zs, nd)) return MCGPU_ERR_CUDA;
        const double* dx = reinterpret_cast<const double*>(d_x);
        const double* dz = reinterpret_cast<const double*>(d_z);
        if (quantized)
            hipLaunchKernelGGL((climate_kernel<double, kQuant5>), dim3(grid_for(count, cus)),
                               dim3(kThreads), 0, stream, d_climate, dx, nullptr, dz, count,
                               (void*)d_out);
        else
            hipLaunchKernelGGL((climate_kernel<double, kRaw5>), dim3(grid_for(count, cus)),
                               dim3(kThreads), 0, stream, d_climate, dx, nullptr, dz, count,
                               (void*)d_out);
        return finish(out, (quantized ? sizeof(int32_t) : sizeof(double)) * 5 * (size_t)count);
    }

    int continentalness(const double* xs, const double* zs, int32_t count, double* out) {
        const size_t nd = sizeof(double) * (size_t)count;
        if (!upload(d_x, xs, nd) || !upload(d_z, zs, nd)) return MCGPU_ERR_CUDA;
        hipLaunchKernelGGL(continentalness_kernel, dim3(grid_for(count, cus)), dim3(kThreads), 0,
                           stream, d_climate, reinterpret_cast<const double*>(d_x),
                           reinterpret_cast<const double*>(d_z), count,
                           reinterpret_cast<double*>(d_out));
        return finish(out, nd);
    }
};

// Filter of candidates of many worlds in one launch (see filter_multi_kernel).
int filter_multi(int32_t nworlds, int tree, const uint64_t* seeds, const int32_t* first,
                 const int32_t* cx, const int32_t* cz, int32_t flags, const Levels& lv,
                 bool large, int32_t* out, float* ms) {
    const int32_t count = first[nworlds];
    const int cus = device_cus();
    TreeArgs targs{};
    TargetLeaves targets{};
    if (!find_targets(kTreeDefs[tree], targets) || !device_tree(tree, targs))
        return MCGPU_ERR_ALLOC;
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
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t ni = sizeof(int32_t) * (size_t)count;
    // the workgroups resident at once (4 per CU at the kernels' register use),
    // at most one per candidate; d_out's last word is the queue counter
    const int32_t grid = std::max(1, std::min(count, cus * 4));
    bool ok = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              hipMalloc(&d_clim, sizeof(Climate) * (size_t)nworlds) == hipSuccess &&
              hipMalloc(&d_seeds, sizeof(int64_t) * (size_t)nworlds) == hipSuccess &&
              hipMalloc(&d_first, sizeof(int32_t) * (size_t)(nworlds + 1)) == hipSuccess &&
              hipMalloc(&d_x, ni) == hipSuccess && hipMalloc(&d_z, ni) == hipSuccess &&
              hipMalloc(&d_out, ni + sizeof(int32_t)) == hipSuccess &&
              hipMemsetAsync(d_out + count, 0, sizeof(int32_t), stream) == hipSuccess &&
              hipMemcpyAsync(d_clim, host.data(), sizeof(Climate) * (size_t)nworlds,
                             hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_seeds, sseed.data(), sizeof(int64_t) * (size_t)nworlds,
                             hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_first, first, sizeof(int32_t) * (size_t)(nworlds + 1),
                             hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_x, cx, ni, hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipMemcpyAsync(d_z, cz, ni, hipMemcpyHostToDevice, stream) == hipSuccess &&
              hipEventRecord(e0, stream) == hipSuccess;
    if (ok) {
        const bool f0 = kTreeDefs[tree].family == 0;
        auto* kernel = lv.n == 0
                           ? (f0 ? filter_multi_kernel<0, false> : filter_multi_kernel<1, false>)
                           : (f0 ? filter_multi_kernel<0, true> : filter_multi_kernel<1, true>);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, stream, d_clim, d_seeds,
                           d_first, nworlds, targs, targets, d_x, d_z, count,
                           reinterpret_cast<unsigned int*>(d_out + count), flags, lv, d_out);
        ok = hipGetLastError() == hipSuccess && hipEventRecord(e1, stream) == hipSuccess &&
             hipMemcpyAsync(out, d_out, ni, hipMemcpyDeviceToHost, stream) == hipSuccess &&
             hipStreamSynchronize(stream) == hipSuccess &&
             hipEventElapsedTime(ms, e0, e1) == hipSuccess;
    }
    for (void* p : {(void*)d_clim, (void*)d_seeds, (void*)d_first, (void*)d_x, (void*)d_z,
                    (void*)d_out})
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

// A new context, or null (the device error, if any, cleared).
Context* new_context(uint64_t seed, bool large, int tree, int32_t capacity, int* rc = nullptr) {
    Context* ctx = new (std::nothrow) Context;
    const int r = ctx ? ctx->init(seed, large, tree, capacity) : MCGPU_ERR_ALLOC;
    if (r != MCGPU_OK) {
        delete ctx;
        ctx = nullptr;
        (void)hipGetLastError();
    }
    if (rc) *rc = r;
    return ctx;
}

int cleared(int rc) {
    if (rc != MCGPU_OK) (void)hipGetLastError();
    return rc;
}

// Runs fn on a context made for one call (the reference's *_fixed_seed calls
// allocate and free per call too).
template <typename Fn>
int one_shot(uint64_t seed, bool large, int tree, int32_t count, Fn fn) {
    int rc = MCGPU_OK;
    Context* ctx = new_context(seed, large, tree, count, &rc);
    if (!ctx) return rc;
    rc = cleared(fn(*ctx));
    delete ctx;
    return rc;
}

// A point count: positive (an int32 count keeps every buffer, at most 48
// bytes per point, far from size_t overflow).
bool count_ok(int32_t count) { return count > 0; }

int tree_20() { return slime::biome::tree_for_version(1206); }

// Flags of the level-range filter: a biome (bit 0 or 1) and optionally bit 2.
bool level_flags_ok(int32_t flags) { return (flags & ~7) == 0 && (flags & 3) != 0; }

// Block heights y_min, y_min + y_step, ... <= y_max: at most kMaxFilterLevels
// heights within +-2^24 blocks (their quart y and sums stay in int32).
constexpr int32_t kMaxFilterLevels = 4096;
bool levels_of(int32_t y_min, int32_t y_max, int32_t y_step, slime::biome::Levels& lv) {
    constexpr int32_t kLimit = 1 << 24;
    if (y_step < 1 || y_min > y_max || y_min < -kLimit || y_max > kLimit) return false;
    const int64_t n = ((int64_t)y_max - y_min) / y_step + 1;
    if (n > kMaxFilterLevels) return false;
    lv = slime::biome::Levels{y_min, y_step, (int32_t)n};
    return true;
}

unsigned blocks_for(int32_t count) {
    return (unsigned)((count + slime::biome::kThreads - 1) / slime::biome::kThreads);
}

}  // namespace

extern "C" {

int mcgpu_api_version(void) { return 1; }

const char* mcgpu_build_name(void) { return "slime_amd biome filter (gfx950, HIP)"; }

void* mcgpu_modern_ctx_create_ex(int32_t version_code, uint64_t world_seed, int32_t flags,
                                 int32_t capacity) {
    const int tree = slime::biome::tree_for_version(version_code);
    if (capacity <= 0 || tree < 0 || (flags & ~1) != 0) return nullptr;
    return new_context(world_seed, (flags & 1) != 0, tree, capacity);
}

void* mcgpu_modern_ctx_create(int32_t version_code, uint64_t world_seed, int32_t capacity) {
    return mcgpu_modern_ctx_create_ex(version_code, world_seed, 0, capacity);
}

void mcgpu_modern_ctx_destroy(void* handle) { delete static_cast<Context*>(handle); }

int mcgpu_modern_ctx_filter_candidates(void* handle, const int32_t* center_chunk_x,
                                       const int32_t* center_chunk_z, int32_t count,
                                       int32_t flags, int32_t* mask_out) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || ctx->tree < 0 || !center_chunk_x || !center_chunk_z || !mask_out || count <= 0 ||
        count > ctx->capacity || (flags & ~3) != 0 || flags == 0)
        return MCGPU_ERR_ARG;
    return cleared(ctx->filter(center_chunk_x, center_chunk_z, count, flags,
                               slime::biome::Levels{0, 0, 0}, mask_out));
}

int mcgpu_modern_ctx_filter_candidates_levels(void* handle, const int32_t* center_chunk_x,
                                              const int32_t* center_chunk_z, int32_t count,
                                              int32_t flags, int32_t y_min, int32_t y_max,
                                              int32_t y_step, int32_t* mask_out) {
    Context* ctx = static_cast<Context*>(handle);
    slime::biome::Levels lv{};
    if (!ctx || ctx->tree < 0 || !center_chunk_x || !center_chunk_z || !mask_out || count <= 0 ||
        count > ctx->capacity || !level_flags_ok(flags) || !levels_of(y_min, y_max, y_step, lv))
        return MCGPU_ERR_ARG;
    return cleared(ctx->filter(center_chunk_x, center_chunk_z, count, flags, lv, mask_out));
}

int mcgpu_modern_ctx_query(void* handle, const int32_t* xs, const int32_t* ys, const int32_t* zs,
                           int32_t count, int32_t* biome_out) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || ctx->tree < 0 || !xs || !ys || !zs || !biome_out || count <= 0 ||
        count > ctx->capacity)
        return MCGPU_ERR_ARG;
    return cleared(ctx->query(xs, ys, zs, count, biome_out));
}

int32_t mcgpu_modern_recommended_filter_batch(void) {
    return slime::biome::device_cus() * 16;  // candidates per call that keep every CU busy
}

int mcgpu_modern_ctx_columns(void* handle, const int32_t* xs, const int32_t* zs, int32_t count,
                             int32_t* out6) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || !xs || !zs || !out6 || count <= 0 || count > ctx->capacity) return MCGPU_ERR_ARG;
    return cleared(ctx->columns(xs, zs, count, out6));
}

namespace {

int filter_multi_checked(int32_t nworlds, int32_t version_code, const uint64_t* world_seeds,
                         const int32_t* first, const int32_t* center_chunk_x,
                         const int32_t* center_chunk_z, int32_t flags,
                         const slime::biome::Levels& lv, int32_t large_biomes,
                         int32_t* mask_out) {
    const int tree = slime::biome::tree_for_version(version_code);
    if (nworlds <= 0 || tree < 0 || !world_seeds || !first || !center_chunk_x ||
        !center_chunk_z || !mask_out || first[0] != 0)
        return MCGPU_ERR_ARG;
    for (int32_t w = 0; w < nworlds; ++w)
        if (first[w + 1] < first[w]) return MCGPU_ERR_ARG;
    if (first[nworlds] == 0) return MCGPU_OK;
    return cleared(slime::biome::filter_multi(nworlds, tree, world_seeds, first, center_chunk_x,
                                              center_chunk_z, flags, lv, large_biomes != 0,
                                              mask_out, &g_multi_ms));
}

}  // namespace

int mcgpu_filter_multi(int32_t nworlds, int32_t version_code, const uint64_t* world_seeds,
                       const int32_t* first, const int32_t* center_chunk_x,
                       const int32_t* center_chunk_z, int32_t flags, int32_t large_biomes,
                       int32_t* mask_out) {
    if ((flags & ~3) != 0 || flags == 0) return MCGPU_ERR_ARG;
    return filter_multi_checked(nworlds, version_code, world_seeds, first, center_chunk_x,
                                center_chunk_z, flags, slime::biome::Levels{0, 0, 0},
                                large_biomes, mask_out);
}

int mcgpu_filter_multi_levels(int32_t nworlds, int32_t version_code, const uint64_t* world_seeds,
                              const int32_t* first, const int32_t* center_chunk_x,
                              const int32_t* center_chunk_z, int32_t flags, int32_t large_biomes,
                              int32_t y_min, int32_t y_max, int32_t y_step, int32_t* mask_out) {
    slime::biome::Levels lv{};
    if (!level_flags_ok(flags) || !levels_of(y_min, y_max, y_step, lv)) return MCGPU_ERR_ARG;
    return filter_multi_checked(nworlds, version_code, world_seeds, first, center_chunk_x,
                                center_chunk_z, flags, lv, large_biomes, mask_out);
}

float mcgpu_filter_multi_last_ms(void) { return g_multi_ms; }

float mcgpu_modern_ctx_last_filter_ms(void* handle) {
    Context* ctx = static_cast<Context*>(handle);
    return ctx ? ctx->last_ms : -1.f;
}

// ---- primitives and point queries ----

int mcgpu_xoroshiro_nextlong_batch(const uint64_t* seeds, int32_t count, int32_t steps,
                                   uint64_t* out) {
    if (!seeds || !out || count <= 0 || steps <= 0) return MCGPU_ERR_ARG;
    const size_t n_out = (size_t)count * (size_t)steps;
    if (n_out > ((size_t)1 << 60)) return MCGPU_ERR_ARG;
    uint64_t *d_seeds = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_seeds, sizeof(uint64_t) * (size_t)count) != hipSuccess ||
        hipMalloc(&d_out, sizeof(uint64_t) * n_out) != hipSuccess) {
        if (d_seeds) (void)hipFree(d_seeds);
        (void)hipGetLastError();
        return MCGPU_ERR_ALLOC;
    }
    int rc = MCGPU_OK;
    if (hipMemcpy(d_seeds, seeds, sizeof(uint64_t) * (size_t)count, hipMemcpyHostToDevice) !=
        hipSuccess) {
        rc = MCGPU_ERR_CUDA;
    } else {
        hipLaunchKernelGGL(slime::biome::xoro_kernel, dim3(blocks_for(count)),
                           dim3(slime::biome::kThreads), 0, nullptr, d_seeds, count, steps, d_out);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, d_out, sizeof(uint64_t) * n_out, hipMemcpyDeviceToHost) != hipSuccess)
            rc = MCGPU_ERR_CUDA;
    }
    (void)hipFree(d_out);
    (void)hipFree(d_seeds);
    return cleared(rc);
}

int mcgpu_xperlin_sample_batch(const uint64_t* seeds, const double* xs, const double* ys,
                               const double* zs, int32_t count, double* out) {
    if (!seeds || !xs || !ys || !zs || !out || !count_ok(count)) return MCGPU_ERR_ARG;
    const size_t n8 = sizeof(double) * (size_t)count;
    uint64_t* d_seeds = nullptr;
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_seeds, n8) != hipSuccess || hipMalloc(&d_x, n8) != hipSuccess ||
        hipMalloc(&d_y, n8) != hipSuccess || hipMalloc(&d_z, n8) != hipSuccess ||
        hipMalloc(&d_out, n8) != hipSuccess) {
        for (void* p : {(void*)d_seeds, (void*)d_x, (void*)d_y, (void*)d_z, (void*)d_out})
            if (p) (void)hipFree(p);
        (void)hipGetLastError();
        return MCGPU_ERR_ALLOC;
    }
    int rc = MCGPU_OK;
    if (hipMemcpy(d_seeds, seeds, n8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_x, xs, n8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_y, ys, n8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_z, zs, n8, hipMemcpyHostToDevice) != hipSuccess) {
        rc = MCGPU_ERR_CUDA;
    } else {
        hipLaunchKernelGGL(slime::biome::perlin_batch_kernel, dim3(blocks_for(count)),
                           dim3(slime::biome::kThreads), 0, nullptr, d_seeds, d_x, d_y, d_z,
                           count, d_out);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, d_out, n8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = MCGPU_ERR_CUDA;
    }
    for (void* p : {(void*)d_seeds, (void*)d_x, (void*)d_y, (void*)d_z, (void*)d_out})
        (void)hipFree(p);
    return cleared(rc);
}

int mcgpu_xperlin_sample_fixed_seed(uint64_t seed, const double* xs, const double* ys,
                                    const double* zs, int32_t count, double* out) {
    if (!xs || !ys || !zs || !out || !count_ok(count)) return MCGPU_ERR_ARG;
    const size_t n8 = sizeof(double) * (size_t)count;
    slime::biome::Perlin host;
    std::memset(&host, 0, sizeof(host));
    slime::biome::perlin_init(host, slime::biome::xoro_from_seed(seed));
    slime::biome::Perlin* d_state = nullptr;
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_state, sizeof(host)) != hipSuccess || hipMalloc(&d_x, n8) != hipSuccess ||
        hipMalloc(&d_y, n8) != hipSuccess || hipMalloc(&d_z, n8) != hipSuccess ||
        hipMalloc(&d_out, n8) != hipSuccess) {
        for (void* p : {(void*)d_state, (void*)d_x, (void*)d_y, (void*)d_z, (void*)d_out})
            if (p) (void)hipFree(p);
        (void)hipGetLastError();
        return MCGPU_ERR_ALLOC;
    }
    int rc = MCGPU_OK;
    if (hipMemcpy(d_state, &host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_x, xs, n8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_y, ys, n8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_z, zs, n8, hipMemcpyHostToDevice) != hipSuccess) {
        rc = MCGPU_ERR_CUDA;
    } else {
        const int grid = slime::biome::grid_for(count, slime::biome::device_cus());
        hipLaunchKernelGGL(slime::biome::perlin_fixed_kernel, dim3(grid),
                           dim3(slime::biome::kThreads), 0, nullptr, d_state, d_x, d_y, d_z,
                           count, d_out);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, d_out, n8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = MCGPU_ERR_CUDA;
    }
    for (void* p : {(void*)d_state, (void*)d_x, (void*)d_y, (void*)d_z, (void*)d_out})
        (void)hipFree(p);
    return cleared(rc);
}

int mcgpu_continentalness_noise_fixed_seed(uint64_t world_seed, const double* xs,
                                           const double* zs, int32_t count, double* out) {
    if (!xs || !zs || !out || !count_ok(count)) return MCGPU_ERR_ARG;
    return one_shot(world_seed, false, -1, count,
                    [&](Context& c) { return c.continentalness(xs, zs, count, out); });
}

int mcgpu_climate5_fixed_seed(uint64_t world_seed, const double* xs, const double* zs,
                              int32_t count, double* out5) {
    if (!xs || !zs || !out5 || !count_ok(count)) return MCGPU_ERR_ARG;
    return one_shot(world_seed, false, -1, count,
                    [&](Context& c) { return c.climate_f64(xs, zs, count, false, out5); });
}

int mcgpu_climate5_quant_fixed_seed(uint64_t world_seed, const double* xs, const double* zs,
                                    int32_t count, int32_t* out5) {
    if (!xs || !zs || !out5 || !count_ok(count)) return MCGPU_ERR_ARG;
    return one_shot(world_seed, false, -1, count,
                    [&](Context& c) { return c.climate_f64(xs, zs, count, true, out5); });
}

int mcgpu_climate5_i32_fixed_seed(uint64_t world_seed, const int32_t* xs, const int32_t* zs,
                                  int32_t count, int32_t* out5) {
    if (!xs || !zs || !out5 || !count_ok(count)) return MCGPU_ERR_ARG;
    return one_shot(world_seed, false, -1, count,
                    [&](Context& c) { return c.climate_i32(xs, nullptr, zs, count, out5); });
}

void* mcgpu_climate_ctx_create(uint64_t world_seed, int32_t capacity) {
    if (!count_ok(capacity)) return nullptr;
    return new_context(world_seed, false, -1, capacity);
}

int mcgpu_climate_ctx_query_i32(void* handle, const int32_t* xs, const int32_t* zs,
                                int32_t count, int32_t* out5) {
    Context* ctx = static_cast<Context*>(handle);
    if (!ctx || !xs || !zs || !out5 || count <= 0 || count > ctx->capacity) return MCGPU_ERR_ARG;
    return cleared(ctx->climate_i32(xs, nullptr, zs, count, out5));
}

void mcgpu_climate_ctx_destroy(void* handle) { delete static_cast<Context*>(handle); }

int mcgpu_climate6_i32_fixed_seed_ex(uint64_t world_seed, int32_t large, const int32_t* xs,
                                     const int32_t* ys, const int32_t* zs, int32_t count,
                                     int32_t* out6) {
    if (!xs || !ys || !zs || !out6 || !count_ok(count)) return MCGPU_ERR_ARG;
    return one_shot(world_seed, large != 0, -1, count,
                    [&](Context& c) { return c.climate_i32(xs, ys, zs, count, out6); });
}

int mcgpu_climate6_i32_fixed_seed(uint64_t world_seed, const int32_t* xs, const int32_t* ys,
                                  const int32_t* zs, int32_t count, int32_t* out6) {
    return mcgpu_climate6_i32_fixed_seed_ex(world_seed, 0, xs, ys, zs, count, out6);
}

int mcgpu_biome20_i32_fixed_seed(uint64_t world_seed, const int32_t* xs, const int32_t* ys,
                                 const int32_t* zs, int32_t count, int32_t* biome_out) {
    if (!xs || !ys || !zs || !biome_out || !count_ok(count)) return MCGPU_ERR_ARG;
    return one_shot(world_seed, false, tree_20(), count,
                    [&](Context& c) { return c.query(xs, ys, zs, count, biome_out); });
}

void* mcgpu_biome20_ctx_create(uint64_t world_seed, int32_t capacity) {
    if (!count_ok(capacity)) return nullptr;
    return new_context(world_seed, false, tree_20(), capacity);
}

int mcgpu_biome20_ctx_query(void* handle, const int32_t* xs, const int32_t* ys, const int32_t* zs,
                            int32_t count, int32_t* biome_out) {
    return mcgpu_modern_ctx_query(handle, xs, ys, zs, count, biome_out);
}

void mcgpu_biome20_ctx_destroy(void* handle) { delete static_cast<Context*>(handle); }

}  // extern "C"
