/* rsh_engine.h -- drop-in C ABI for the reference's RSH stage wrappers.
 *
 * Same entry-point names, parameter lists and buffer layouts as the
 * extern "C" block of projects/CUDA-Engine/source/RayTracer.cu:2042-3301
 * (declared again by the caller at projects/CUDA-Engine/source/main.cpp:192-583).
 * Each declaration cites the reference function it replaces.
 *
 * Differences a caller must know (the integration notes show the edits):
 *  - float3/float4 become rsh_float3/rsh_float4, layout-identical POD structs
 *    (12-byte float3 stride, 16-byte aligned float4).
 *  - gfx950 has no texture references; the bind* calls store device pointers.
 *    The four G-buffer binders take a device buffer of W*H float4 (row-major,
 *    pitch W) instead of a cudaArray* from GL interop.
 *  - The compile-time constants LIGHT_SOURCE_MAXIMUM, HIERARCHY_MAXIMUM_DEPTH,
 *    HIERARCHY_SUBDIVISION (Constants.h:15-29) and the global algorithmID
 *    (main.cpp:98) are set with rsh_compat_configure; defaults 1, 2, 8, CRSH.
 *  - Errors: by default a failing call prints the error and exits(1) like
 *    Utility::checkCUDAError (Utility.h:72-81); rsh_compat_set_exit_on_error(0)
 *    turns that into a sticky status read with rsh_compat_status().
 *  - Wrappers that returned counts through host pointers after a blocking
 *    cudaMemcpy keep doing so (RayTracer.cu:2316,2364,2649,2720,2832).
 *  - Auxiliary buffers (head flags, scan arrays, skeleton) are scratch: their
 *    contents after a call are unspecified, exactly as in the reference where
 *    the caller never reads them.
 */
#ifndef RSH_ENGINE_H_
#define RSH_ENGINE_H_

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsh_float2 { float x, y; } rsh_float2;
typedef struct rsh_float3 { float x, y, z; } rsh_float3;
typedef struct __attribute__((aligned(16))) rsh_float4 { float x, y, z, w; } rsh_float4;

/* ---------------------------------------------------------------- setup -- */
/* Constants.h:15 LIGHT_SOURCE_MAXIMUM, :28 HIERARCHY_MAXIMUM_DEPTH,
 * :29 HIERARCHY_SUBDIVISION, main.cpp:98 algorithmID (0 CRSH, 1 RAH). */
void rsh_compat_configure(unsigned int lightMaximum, unsigned int hierarchyDepth,
                          unsigned int hierarchySubdivision, int algorithmID);
void rsh_compat_set_stream(void* hipStream);
void rsh_compat_set_exit_on_error(int exitOnError);
int rsh_compat_status(void);               /* 0 = OK, else an rsh_status code */
const char* rsh_compat_error(void);
void rsh_compat_reset_status(void);

/* RayTracer.cu:2046 TriangleUpdateWrapper */
void TriangleUpdateWrapper(float* modelMatricesArray, float* normalMatricesArray,
                           unsigned int triangleTotal, rsh_float4* trianglePositionsArray,
                           rsh_float4* triangleNormalsArray);
/* RayTracer.cu:2071 BoundingSphereUpdateWrapper */
void BoundingSphereUpdateWrapper(float* translationMatricesArray, float* scaleMatricesArray,
                                 const unsigned int boundingSphereTotal,
                                 rsh_float3* boundingSphereArray);
/* RayTracer.cu:2093 MemoryPreparationWrapper (sizes the library's scan and sort scratch) */
void MemoryPreparationWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                              const unsigned int hierarchyHitMemoryTotal,
                              unsigned int* headFlagsArray, unsigned int* scanArray,
                              unsigned int* chunkIndexKeysArray, unsigned int* chunkIndexValuesArray,
                              unsigned int* sortedChunkIndexKeysArray,
                              unsigned int* sortedChunkIndexValuesArray);
/* RayTracer.cu:2210 ShadowRayCreationWrapper */
void ShadowRayCreationWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                              const unsigned int lightTotal, rsh_float3* rayArray,
                              unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                              unsigned int* rayIndexValuesArray);
/* No counterpart wrapper in the reference: launches its CreateRefractionRays
   (RayTracer.cu:793-842, which no reference code calls) with the arguments of
   ReflectionRayCreationWrapper; refraction index 1/1.52, the reference's
   CreateRefractionRayIndex keys (RayTracer.cu:200-227). */
void RefractionRayCreationWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                                  const rsh_float3 cameraPosition, rsh_float3* rayArray,
                                  unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                                  unsigned int* rayIndexValuesArray);
/* RayTracer.cu:2243 ReflectionRayCreationWrapper */
void ReflectionRayCreationWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                                  const rsh_float3 cameraPosition, rsh_float3* rayArray,
                                  unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                                  unsigned int* rayIndexValuesArray);
/* RayTracer.cu:2276 RayTrimmingWrapper */
void RayTrimmingWrapper(unsigned int* rayIndexKeysArray, unsigned int* rayIndexValuesArray,
                        const unsigned int windowWidth, const unsigned int windowHeight,
                        const unsigned int lightTotal, unsigned int* headFlagsArray,
                        unsigned int* scanArray, unsigned int* trimmedRayIndexKeysArray,
                        unsigned int* trimmedRayIndexValuesArray, unsigned int* rayTotal);
/* RayTracer.cu:2326 RayCompressionWrapper */
void RayCompressionWrapper(unsigned int* trimmedRayIndexKeysArray,
                           unsigned int* trimmedRayIndexValuesArray, const unsigned int rayTotal,
                           unsigned int* headFlagsArray, unsigned int* scanArray,
                           unsigned int* chunkBasesArray, unsigned int* chunkSizesArray,
                           unsigned int* chunkIndexKeysArray, unsigned int* chunkIndexValuesArray,
                           unsigned int* chunkTotal);
/* No counterpart in the reference: the sorting stage on the low `bits` bits
   of the keys (stable LSD radix sort, 8- or 9-bit digits), as the engine sorts
   a frame's shadow and reflection passes (27 of their 29 key bits vary). */
void rsh_sort_pairs(const unsigned int* keys, const unsigned int* values, const unsigned int n,
                    const unsigned int bits, unsigned int* out_keys, unsigned int* out_values);
/* No counterpart in the reference.  For callers that cannot allocate (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc): rsh_sort_pairs with the sort's scratch in a file-scope device
   array (17 MiB per GPU, held until the process ends) instead of library
   allocations.  n <= 2^21 pairs (2,097,152: a 1920x1080 frame's slots); more
   returns RSH_CAPACITY, bits outside 1..32 or a null pointer RSH_INVALID_ARG.
   Asynchronous on `stream` (a hipStream_t; NULL = the null stream), no
   synchronisation, no allocation.  Every call on one GPU shares that GPU's
   array: calls that overlap on one GPU race.  The first call on a device looks
   the array up (outside a graph capture).  Returns an rsh_status code
   (include/rsh.h): 0 OK, 1 invalid argument, 2 capacity, 3 HIP error. */
int rsh_sort_pairs_noalloc(const uint32_t* keys, const uint32_t* values, uint32_t n, uint32_t bits,
                           uint32_t* out_keys, uint32_t* out_values, void* stream);
/* RayTracer.cu:2385 RaySortingWrapper (CUB SortPairs replaced by the library's radix sort) */
void RaySortingWrapper(unsigned int* chunkIndexKeysArray, unsigned int* chunkIndexValuesArray,
                       const unsigned int chunkTotal, unsigned int* sortedChunkIndexKeysArray,
                       unsigned int* sortedChunkIndexValuesArray);
/* RayTracer.cu:2411 RayDecompressionWrapper */
void RayDecompressionWrapper(unsigned int* chunkBasesArray, unsigned int* chunkSizesArray,
                             unsigned int* sortedChunkIndexKeysArray,
                             unsigned int* sortedChunkIndexValuesArray,
                             unsigned int* trimmedRayIndexKeysArray,
                             unsigned int* trimmedRayIndexValuesArray, unsigned int* skeletonArray,
                             unsigned int* scanArray, const unsigned int chunkTotal,
                             unsigned int* sortedRayIndexKeysArray,
                             unsigned int* sortedRayIndexValuesArray);
/* RayTracer.cu:2469 HierarchyCreationWrapper */
void HierarchyCreationWrapper(rsh_float3* rayArray, unsigned int* sortedRayIndexKeysArray,
                              unsigned int* sortedRayIndexValuesArray, const unsigned int rayTotal,
                              const float initialRadius, const float initialSpread,
                              rsh_float4* hierarchyArray);
/* RayTracer.cu:2537 HierarchyTraversalWarmUpWrapper */
void HierarchyTraversalWarmUpWrapper(rsh_float4* hierarchyArray, rsh_float3* boundingSphereArray,
                                     rsh_float4* trianglePositionsArray,
                                     const unsigned int boundingSphereTotal,
                                     const unsigned int rayTotal, const unsigned int triangleTotal,
                                     const unsigned int triangleOffset, unsigned int* headFlagsArray,
                                     unsigned int* scanArray, unsigned int* hierarchyHitsArray,
                                     unsigned int* trimmedHierarchyHitsArray,
                                     unsigned int* hierarchyHitTotal,
                                     unsigned int* hierarchyHitMemoryTotal);
/* RayTracer.cu:2741 HierarchyTraversalWrapper */
void HierarchyTraversalWrapper(rsh_float4* hierarchyArray, rsh_float4* trianglePositionsArray,
                               const unsigned int rayTotal, const unsigned int triangleTotal,
                               const unsigned int triangleOffset, unsigned int* headFlagsArray,
                               unsigned int* scanArray, unsigned int* hierarchyHitsArray,
                               unsigned int* trimmedHierarchyHitsArray,
                               unsigned int* hierarchyHitTotal,
                               unsigned int* hierarchyHitMemoryTotal);
/* RayTracer.cu:2854 ShadowRayPreparationWrapper */
void ShadowRayPreparationWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                                 const unsigned int lightTotal, unsigned int* shadowFlagsArray);
/* RayTracer.cu:2874 ShadowRayIntersectionWrapper */
void ShadowRayIntersectionWrapper(rsh_float3* rayArray, unsigned int* sortedRayIndexKeysArray,
                                  unsigned int* sortedRayIndexValuesArray,
                                  unsigned int* hierarchyHitsArray,
                                  rsh_float4* trianglePositionsArray, const unsigned int hitTotal,
                                  const unsigned int rayTotal, const unsigned int triangleOffset,
                                  const unsigned int windowWidth, const unsigned int windowHeight,
                                  const rsh_float3 cameraPosition, unsigned int* shadowFlagsArray);
/* RayTracer.cu:2927 ShadowRayColoringWrapper */
void ShadowRayColoringWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                              const unsigned int lightTotal, const rsh_float3 cameraPosition,
                              unsigned int* shadowFlagsArray, unsigned int* pixelBufferObject);
/* RayTracer.cu:2959 ReflectionRayPreparationWrapper */
void ReflectionRayPreparationWrapper(const unsigned int windowWidth, const unsigned int windowHeight,
                                     unsigned int* intersectionTimeArray);
/* RayTracer.cu:2977 ReflectionRayIntersectionWrapper */
void ReflectionRayIntersectionWrapper(rsh_float3* rayArray, unsigned int* sortedRayIndexKeysArray,
                                      unsigned int* sortedRayIndexValuesArray,
                                      unsigned int* hierarchyHitsArray,
                                      rsh_float4* trianglePositionsArray,
                                      const unsigned int hitTotal, const unsigned int rayTotal,
                                      const unsigned int triangleOffset,
                                      const unsigned int windowWidth,
                                      const unsigned int windowHeight,
                                      const unsigned int lightTotal,
                                      const rsh_float3 cameraPosition,
                                      unsigned int* intersectionTimeArray);
/* RayTracer.cu:3033 ReflectionRayColoringWrapper */
void ReflectionRayColoringWrapper(rsh_float4* trianglePositionsArray,
                                  rsh_float4* triangleNormalsArray, const unsigned int windowWidth,
                                  const unsigned int windowHeight, const unsigned int lightTotal,
                                  const rsh_float3 cameraPosition, const bool createRays,
                                  unsigned int* intersectionTimeArray, rsh_float3* rayArray,
                                  unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                                  unsigned int* rayIndexValuesArray,
                                  unsigned int* pixelBufferObject);

/* ------------------------------------------------- binders (RayTracer.cu:3123-3301) */
void bindDiffuseTextureArray(const rsh_float4* diffuseBuffer);           /* :3123 */
void bindSpecularTextureArray(const rsh_float4* specularBuffer);         /* :3134 */
void bindFragmentPositionArray(const rsh_float4* fragmentPositionBuffer); /* :3145 */
void bindFragmentNormalArray(const rsh_float4* fragmentNormalBuffer);    /* :3156 */
void bindTrianglePositions(float* cudaDevicePointer, unsigned int triangleTotal);          /* :3168 */
void bindTriangleNormals(float* cudaDevicePointer, unsigned int triangleTotal);            /* :3180 */
void bindTriangleTextureCoordinates(float* cudaDevicePointer, unsigned int triangleTotal); /* :3192 */
void bindTriangleObjectIDs(float* cudaDevicePointer, unsigned int triangleTotal);          /* :3204 */
void bindTriangleMaterialIDs(float* cudaDevicePointer, unsigned int triangleTotal);        /* :3216 */
void bindMaterialDiffuseProperties(float* cudaDevicePointer, unsigned int materialTotal);  /* :3229 */
void bindMaterialSpecularProperties(float* cudaDevicePointer, unsigned int materialTotal); /* :3241 */
void bindBoundingSpheres(float* cudaDevicePointer, unsigned int boundingSphereTotal);      /* :3254 */
void bindLightPositions(float* cudaDevicePointer, unsigned int lightTotal);                /* :3267 */
void bindLightColors(float* cudaDevicePointer, unsigned int lightTotal);                   /* :3279 */
void bindLightIntensities(float* cudaDevicePointer, unsigned int lightTotal);              /* :3291 */

#ifdef __cplusplus
}
#endif

#endif /* RSH_ENGINE_H_ */
