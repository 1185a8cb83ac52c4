/* Biome filter of the slime-cluster search: the reference's gpu_noise C ABI
 * (gpu_noise/mcgpu.h, MinecraftGPUNoise.cu:1076-2018) on MI355X.
 *
 * A context holds one world's overworld biome source (Minecraft's multi-noise
 * sampler and the parameter tree of its version) on the current HIP device. Return
 * codes: MCGPU_OK, MCGPU_ERR_ARG for invalid arguments, MCGPU_ERR_CUDA for a
 * device error (the name is the reference's). */
#ifndef SLIME_MCGPU_H
#define SLIME_MCGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MCGPU_OK = 0, MCGPU_ERR_ARG = -1, MCGPU_ERR_CUDA = -2, MCGPU_ERR_ALLOC = -3 };

/* MinecraftGPUNoise.cu:1076-1079: the ABI version, 1. */
int mcgpu_api_version(void);
/* MinecraftGPUNoise.cu:1102-1105: a descriptive name of the build. */
const char* mcgpu_build_name(void);

/* MinecraftGPUNoise.cu:1837-1880: a context for world_seed with buffers for
 * `capacity` candidates or points; NULL on failure. The version code selects
 * the parameter tree as MinecraftGPUNoise.cu:1785-1820 does (>= 2602: 26.2,
 * >= 1215, >= 1214, >= 1206 (1.20), >= 1194, >= 1192, >= 1180 (1.18); below
 * 1180: NULL). */
void* mcgpu_modern_ctx_create(int32_t version_code, uint64_t world_seed, int32_t capacity);
/* MinecraftGPUNoise.cu:1950-1966: flags bit 0 = large biomes. */
void* mcgpu_modern_ctx_create_ex(int32_t version_code, uint64_t world_seed, int32_t flags,
                                 int32_t capacity);
/* MinecraftGPUNoise.cu:1935-1948. */
void mcgpu_modern_ctx_destroy(void* ctx);

/* MinecraftGPUNoise.cu:1910-1933 (kernel :978-1074): for each candidate
 * center chunk, mask_out gets bit 0 if some slime chunk of its 221-chunk disk
 * has Deep Dark at a quart of y -16..0 and bit 1 if one has Mushroom Fields at
 * quart y 16, for the bits requested in flags (1, 2 or 3). */
int mcgpu_modern_ctx_filter_candidates(void* ctx, const int32_t* center_chunk_x,
                                       const int32_t* center_chunk_z, int32_t count,
                                       int32_t flags, int32_t* mask_out);

/* Extension with no counterpart in the reference: the filter over a range of
 * heights, for the wider sweep (y -64..320 step 4) next to the reference's
 * fixed set. Each biome requested in flags (bit 0 Deep Dark, bit 1 Mushroom
 * Fields) is looked for at block heights y_min, y_min + y_step, ... <= y_max,
 * sampled at quart y floor(y / 4), in the 16 quart columns of each sampled
 * chunk; flags bit 2 samples every chunk of the 221-chunk disk instead of its
 * slime chunks. MCGPU_ERR_ARG unless y_step >= 1, y_min <= y_max, both within
 * +-2^24 and at most 4096 heights. With heights -64..0 step 4 (flags 1) or 64
 * alone (flags 2) it gives the masks of mcgpu_modern_ctx_filter_candidates. */
enum { MCGPU_FILTER_ALL_CHUNKS = 4 };
int mcgpu_modern_ctx_filter_candidates_levels(void* ctx, const int32_t* center_chunk_x,
                                              const int32_t* center_chunk_z, int32_t count,
                                              int32_t flags, int32_t y_min, int32_t y_max,
                                              int32_t y_step, int32_t* mask_out);

/* MinecraftGPUNoise.cu:1882-1908: biome ids at quart coordinates. */
int mcgpu_modern_ctx_query(void* ctx, const int32_t* xs, const int32_t* ys, const int32_t* zs,
                           int32_t count, int32_t* biome_out);

/* MinecraftGPUNoise.cu:1081-1100: a batch size for filter calls. */
int32_t mcgpu_modern_recommended_filter_batch(void);

/* Extension with no counterpart in the reference: the y-independent climate of
 * quart columns, six int32 per point (quantized temperature, humidity,
 * continentalness, erosion, weirdness, and the bits of the float depth-spline
 * offset), for checking the noise stage on its own. */
int mcgpu_modern_ctx_columns(void* ctx, const int32_t* xs, const int32_t* zs, int32_t count,
                             int32_t* out6);

/* Extension with no counterpart in the reference: device time (ms, HIP events)
 * of the last filter call of the context. */
float mcgpu_modern_ctx_last_filter_ms(void* ctx);

/* Extension with no counterpart in the reference: the filter of candidates of
 * many worlds in one launch (the batched search's top sites per seed). World w
 * has seed world_seeds[w] and candidates [first[w], first[w+1]) of the arrays
 * (first[0] == 0); version_code, flags and mask_out as for a context and
 * mcgpu_modern_ctx_filter_candidates; large_biomes != 0 selects the
 * large-biome noises. */
int mcgpu_filter_multi(int32_t nworlds, int32_t version_code, const uint64_t* world_seeds,
                       const int32_t* first, const int32_t* center_chunk_x,
                       const int32_t* center_chunk_z, int32_t flags, int32_t large_biomes,
                       int32_t* mask_out);
/* Extension with no counterpart in the reference: mcgpu_filter_multi over a
 * range of heights, flags and heights as for
 * mcgpu_modern_ctx_filter_candidates_levels. */
int mcgpu_filter_multi_levels(int32_t nworlds, int32_t version_code, const uint64_t* world_seeds,
                              const int32_t* first, const int32_t* center_chunk_x,
                              const int32_t* center_chunk_z, int32_t flags, int32_t large_biomes,
                              int32_t y_min, int32_t y_max, int32_t y_step, int32_t* mask_out);
/* Extension with no counterpart in the reference: device time (ms, HIP events)
 * of the last mcgpu_filter_multi or mcgpu_filter_multi_levels launch. */
float mcgpu_filter_multi_last_ms(void);

/* ---- primitives and point queries (MinecraftGPUNoise.cu:1107-1774, 1988-2018)
 * Each returns MCGPU_OK, MCGPU_ERR_ARG (null pointer or count <= 0),
 * MCGPU_ERR_ALLOC or MCGPU_ERR_CUDA. Quantized climate values are
 * (int32_t)(10000.0F * (float)v). */

/* :1107-1151: `steps` xoroshiro128++ outputs of each seed (Minecraft's
 * seeding), out[i * steps + j]. */
int mcgpu_xoroshiro_nextlong_batch(const uint64_t* seeds, int32_t count, int32_t steps,
                                   uint64_t* out);
/* :1153-1205: improved Perlin noise seeded per query from seeds[i]. */
int mcgpu_xperlin_sample_batch(const uint64_t* seeds, const double* xs, const double* ys,
                               const double* zs, int32_t count, double* out);
/* :1207-1257: one Perlin noise (from seed) at every query point. */
int mcgpu_xperlin_sample_fixed_seed(uint64_t seed, const double* xs, const double* ys,
                                    const double* zs, int32_t count, double* out);
/* :1259-1305: the continentalness double Perlin noise at (x, 0, z), unshifted,
 * unquantized. */
int mcgpu_continentalness_noise_fixed_seed(uint64_t world_seed, const double* xs,
                                           const double* zs, int32_t count, double* out);
/* :1307-1354: five climate noises (temperature, humidity, continentalness,
 * erosion, weirdness) at the shifted point of (x, z), as doubles. */
int mcgpu_climate5_fixed_seed(uint64_t world_seed, const double* xs, const double* zs,
                              int32_t count, double* out5);
/* :1356-1403: the same, quantized. */
int mcgpu_climate5_quant_fixed_seed(uint64_t world_seed, const double* xs, const double* zs,
                                    int32_t count, int32_t* out5);
/* :1405-1459: the same at int32 quart coordinates. */
int mcgpu_climate5_i32_fixed_seed(uint64_t world_seed, const int32_t* xs, const int32_t* zs,
                                  int32_t count, int32_t* out5);
/* :1461-1542: a climate context (NULL on failure) and its quantized queries. */
void* mcgpu_climate_ctx_create(uint64_t world_seed, int32_t capacity);
int mcgpu_climate_ctx_query_i32(void* ctx, const int32_t* xs, const int32_t* zs, int32_t count,
                                int32_t* out5);
void mcgpu_climate_ctx_destroy(void* ctx);
/* :1544-1595: six quantized values (temperature, humidity, continentalness,
 * erosion, depth at quart y, weirdness). */
int mcgpu_climate6_i32_fixed_seed(uint64_t world_seed, const int32_t* xs, const int32_t* ys,
                                  const int32_t* zs, int32_t count, int32_t* out6);
/* :1988-2018: the same with large biomes when large != 0. */
int mcgpu_climate6_i32_fixed_seed_ex(uint64_t world_seed, int32_t large, const int32_t* xs,
                                     const int32_t* ys, const int32_t* zs, int32_t count,
                                     int32_t* out6);
/* :1597-1774: biome ids with the 1.20 tree, one call or a context. */
int mcgpu_biome20_i32_fixed_seed(uint64_t world_seed, const int32_t* xs, const int32_t* ys,
                                 const int32_t* zs, int32_t count, int32_t* biome_out);
void* mcgpu_biome20_ctx_create(uint64_t world_seed, int32_t capacity);
int mcgpu_biome20_ctx_query(void* ctx, const int32_t* xs, const int32_t* ys, const int32_t* zs,
                            int32_t count, int32_t* biome_out);
void mcgpu_biome20_ctx_destroy(void* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SLIME_MCGPU_H */
