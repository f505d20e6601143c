/* rsh_ctx.h -- the reference's stage wrappers as a per-context, per-stream C API.
 *
 * The extern "C" wrappers of projects/CUDA-Engine/source/RayTracer.cu:2042-3301
 * keep their state in globals (texture references for the G-buffer and scene
 * arrays, algorithmID, CUB temporary storage) and run on the legacy default
 * stream.  Here every stage is a function of an explicit context (its bound
 * G-buffer and scene pointers, constants and scratch) and an explicit HIP
 * stream, and returns an rsh_status instead of exiting.  The buffer arguments
 * are the reference wrapper's, in the same order and layout; the drop-in
 * wrappers of include/rsh_engine.h are these functions on one process-wide
 * context and the stream set with rsh_compat_set_stream.
 *
 * Threading: calls on different contexts are independent and may run on
 * different host threads at once.  One context must not be used from two
 * threads at once, and its calls share its scratch, so two calls of one
 * context must be ordered (the same stream, or an event between streams).
 *
 * Counts: the stages that end in a count (ray, chunk and pair totals) write it
 * to a device word `...Device` when one is given, in stream order, and to the
 * host pointer only when that is given (the call then waits for the stream, as
 * the reference's blocking cudaMemcpy does, RayTracer.cu:2316,2364,2649,2720,
 * 2832).  The stages that take a count accept a device word `...Device`
 * instead: the host argument is then the capacity the launch covers and the
 * kernels read the count from device memory.  The chain trimming ->
 * compression -> sorting -> decompression -> hierarchy creation so runs with
 * no host synchronisation and can be captured in a HIP graph.  The traversal
 * stages size each level's launch from the previous level's pair count, so
 * they take host counts and wait, as in the reference.
 *
 * Every function returns RSH_OK (0) or an rsh_status code (include/rsh.h);
 * rsh_ctx_last_error gives the message of the context's last failure.
 * `stream` is a hipStream_t (NULL = the null stream).
 */
#ifndef RSH_CTX_H_
#define RSH_CTX_H_

#include "rsh.h"
#include "rsh_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsh_ctx rsh_ctx;

/* A context on HIP device `device`, with the reference's default constants
   (LIGHT_SOURCE_MAXIMUM 1, HIERARCHY_MAXIMUM_DEPTH 2, HIERARCHY_SUBDIVISION 8,
   CRSH).  Every later call of the context makes `device` current on the
   calling thread. */
rsh_status rsh_ctx_create(int device, rsh_ctx** out);
void rsh_ctx_destroy(rsh_ctx* ctx);
/* Constants.h:15 LIGHT_SOURCE_MAXIMUM, :28 HIERARCHY_MAXIMUM_DEPTH,
   :29 HIERARCHY_SUBDIVISION and main.cpp:98 algorithmID (0 CRSH, 1 RAH) as
   context fields. */
rsh_status rsh_ctx_configure(rsh_ctx* ctx, unsigned int lightMaximum, unsigned int hierarchyDepth,
                             unsigned int hierarchySubdivision, int algorithmID);
const char* rsh_ctx_last_error(const rsh_ctx* ctx);

/* RayTracer.cu:3123-3156 bind{Diffuse,Specular}TextureArray and
   bindFragment{Position,Normal}Array: the G-buffer, four device buffers of
   width x height float4 (row-major).  Any may be NULL (left unbound). */
rsh_status rsh_set_gbuffer(rsh_ctx* ctx, const rsh_float4* diffuse, const rsh_float4* specular,
                           const rsh_float4* position, const rsh_float4* normal, unsigned int width,
                           unsigned int height);
/* RayTracer.cu:3168-3301, the other bind* calls: triangle positions, normals,
   object and material ids (3 per triangle, T triangles), materials (M),
   object bounding spheres (K, as bindBoundingSpheres) and lights (L).  The
   texture coordinates are not used by any stage and are not taken. */
rsh_status rsh_set_scene(rsh_ctx* ctx, const float* trianglePositions, const float* triangleNormals,
                         const float* triangleObjectIDs, const float* triangleMaterialIDs,
                         unsigned int triangleTotal, const float* materialDiffuse, const float* materialSpecular,
                         unsigned int materialTotal, const float* boundingSpheres, unsigned int boundingSphereTotal,
                         const float* lightPositions, const float* lightColors, const float* lightIntensities,
                         unsigned int lightTotal);

/* RayTracer.cu:2046 TriangleUpdateWrapper */
rsh_status rsh_triangle_update(rsh_ctx* ctx, void* stream, float* modelMatricesArray, float* normalMatricesArray,
                               unsigned int triangleTotal, rsh_float4* trianglePositionsArray,
                               rsh_float4* triangleNormalsArray);
/* RayTracer.cu:2071 BoundingSphereUpdateWrapper */
rsh_status rsh_bounding_sphere_update(rsh_ctx* ctx, void* stream, float* translationMatricesArray,
                                      float* scaleMatricesArray, unsigned int boundingSphereTotal,
                                      rsh_float3* boundingSphereArray);
/* RayTracer.cu:2093 MemoryPreparationWrapper: reserves the context's scratch
   for windowWidth x windowHeight x LIGHT_SOURCE_MAXIMUM slots or
   hierarchyHitMemoryTotal pairs, whichever is larger (the buffer arguments
   are the caller's and are not touched). */
rsh_status rsh_memory_preparation(rsh_ctx* ctx, void* stream, unsigned int windowWidth, unsigned int windowHeight,
                                  unsigned int hierarchyHitMemoryTotal);
/* RayTracer.cu:2210 ShadowRayCreationWrapper */
rsh_status rsh_shadow_ray_creation(rsh_ctx* ctx, void* stream, unsigned int windowWidth, unsigned int windowHeight,
                                   unsigned int lightTotal, rsh_float3* rayArray, unsigned int* headFlagsArray,
                                   unsigned int* rayIndexKeysArray, unsigned int* rayIndexValuesArray);
/* RayTracer.cu:2243 ReflectionRayCreationWrapper */
rsh_status rsh_reflection_ray_creation(rsh_ctx* ctx, void* stream, unsigned int windowWidth,
                                       unsigned int windowHeight, rsh_float3 cameraPosition, rsh_float3* rayArray,
                                       unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                                       unsigned int* rayIndexValuesArray);
/* RefractionRayCreationWrapper of include/rsh_engine.h (the reference's
   never-launched CreateRefractionRays, RayTracer.cu:793-842) */
rsh_status rsh_refraction_ray_creation(rsh_ctx* ctx, void* stream, unsigned int windowWidth,
                                       unsigned int windowHeight, rsh_float3 cameraPosition, rsh_float3* rayArray,
                                       unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                                       unsigned int* rayIndexValuesArray);
/* RayTracer.cu:2276 RayTrimmingWrapper; the ray total to rayTotal (host, may
   be NULL) and rayTotalDevice (device, may be NULL). */
rsh_status rsh_ray_trimming(rsh_ctx* ctx, void* stream, unsigned int* rayIndexKeysArray,
                            unsigned int* rayIndexValuesArray, unsigned int windowWidth, unsigned int windowHeight,
                            unsigned int lightTotal, unsigned int* headFlagsArray, unsigned int* scanArray,
                            unsigned int* trimmedRayIndexKeysArray, unsigned int* trimmedRayIndexValuesArray,
                            unsigned int* rayTotal, unsigned int* rayTotalDevice);
/* RayTracer.cu:2326 RayCompressionWrapper; rayTotalDevice (may be NULL): the
   ray total in device memory, rayTotal then bounds it.  The chunk total to
   chunkTotal (host, may be NULL) and chunkTotalDevice (device, may be NULL). */
rsh_status rsh_ray_compression(rsh_ctx* ctx, void* stream, unsigned int* trimmedRayIndexKeysArray,
                               unsigned int* trimmedRayIndexValuesArray, unsigned int rayTotal,
                               const unsigned int* rayTotalDevice, unsigned int* headFlagsArray,
                               unsigned int* scanArray, unsigned int* chunkBasesArray, unsigned int* chunkSizesArray,
                               unsigned int* chunkIndexKeysArray, unsigned int* chunkIndexValuesArray,
                               unsigned int* chunkTotal, unsigned int* chunkTotalDevice);
/* RayTracer.cu:2385 RaySortingWrapper; chunkTotalDevice (may be NULL): the
   chunk total in device memory, chunkTotal then bounds it (outputs must not
   alias inputs in that case). */
rsh_status rsh_ray_sorting(rsh_ctx* ctx, void* stream, unsigned int* chunkIndexKeysArray,
                           unsigned int* chunkIndexValuesArray, unsigned int chunkTotal,
                           const unsigned int* chunkTotalDevice, unsigned int* sortedChunkIndexKeysArray,
                           unsigned int* sortedChunkIndexValuesArray);
/* rsh_sort_pairs of include/rsh_engine.h on a context and stream, with the
   optional device count; no synchronisation. */
rsh_status rsh_ctx_sort_pairs(rsh_ctx* ctx, void* stream, const unsigned int* keys, const unsigned int* values,
                              unsigned int n, const unsigned int* nDevice, unsigned int bits, unsigned int* out_keys,
                              unsigned int* out_values);
/* RayTracer.cu:2411 RayDecompressionWrapper; chunkTotalDevice as in
   rsh_ray_sorting.  The ray total is read on the device (no synchronisation
   either way). */
rsh_status rsh_ray_decompression(rsh_ctx* ctx, void* stream, unsigned int* chunkBasesArray,
                                 unsigned int* chunkSizesArray, unsigned int* sortedChunkIndexKeysArray,
                                 unsigned int* sortedChunkIndexValuesArray, unsigned int* trimmedRayIndexKeysArray,
                                 unsigned int* trimmedRayIndexValuesArray, unsigned int* skeletonArray,
                                 unsigned int* scanArray, unsigned int chunkTotal,
                                 const unsigned int* chunkTotalDevice, unsigned int* sortedRayIndexKeysArray,
                                 unsigned int* sortedRayIndexValuesArray);
/* RayTracer.cu:2469 HierarchyCreationWrapper; rayTotalDevice (may be NULL):
   the ray total in device memory, rayTotal then bounds it (hierarchyArray
   holds the levels of rayTotal rays; the levels of the device total are laid
   out from it as for a host total). */
rsh_status rsh_hierarchy_creation(rsh_ctx* ctx, void* stream, rsh_float3* rayArray,
                                  unsigned int* sortedRayIndexKeysArray, unsigned int* sortedRayIndexValuesArray,
                                  unsigned int rayTotal, const unsigned int* rayTotalDevice, float initialRadius,
                                  float initialSpread, rsh_float4* hierarchyArray);
/* RayTracer.cu:2537 HierarchyTraversalWarmUpWrapper (host counts; waits) */
rsh_status rsh_hierarchy_traversal_warmup(rsh_ctx* ctx, void* stream, rsh_float4* hierarchyArray,
                                          rsh_float3* boundingSphereArray, rsh_float4* trianglePositionsArray,
                                          unsigned int boundingSphereTotal, unsigned int rayTotal,
                                          unsigned int triangleTotal, unsigned int triangleOffset,
                                          unsigned int* headFlagsArray, unsigned int* scanArray,
                                          unsigned int* hierarchyHitsArray, unsigned int* trimmedHierarchyHitsArray,
                                          unsigned int* hierarchyHitTotal, unsigned int* hierarchyHitMemoryTotal);
/* RayTracer.cu:2741 HierarchyTraversalWrapper (host counts; waits) */
rsh_status rsh_hierarchy_traversal(rsh_ctx* ctx, void* stream, rsh_float4* hierarchyArray,
                                   rsh_float4* trianglePositionsArray, unsigned int rayTotal,
                                   unsigned int triangleTotal, unsigned int triangleOffset,
                                   unsigned int* headFlagsArray, unsigned int* scanArray,
                                   unsigned int* hierarchyHitsArray, unsigned int* trimmedHierarchyHitsArray,
                                   unsigned int* hierarchyHitTotal, unsigned int* hierarchyHitMemoryTotal);
/* RayTracer.cu:2854 ShadowRayPreparationWrapper */
rsh_status rsh_shadow_ray_preparation(rsh_ctx* ctx, void* stream, unsigned int windowWidth, unsigned int windowHeight,
                                      unsigned int lightTotal, unsigned int* shadowFlagsArray);
/* RayTracer.cu:2874 ShadowRayIntersectionWrapper */
rsh_status rsh_shadow_ray_intersection(rsh_ctx* ctx, void* stream, rsh_float3* rayArray,
                                       unsigned int* sortedRayIndexKeysArray, unsigned int* sortedRayIndexValuesArray,
                                       unsigned int* hierarchyHitsArray, rsh_float4* trianglePositionsArray,
                                       unsigned int hitTotal, unsigned int rayTotal, unsigned int triangleOffset,
                                       unsigned int windowWidth, unsigned int windowHeight,
                                       rsh_float3 cameraPosition, unsigned int* shadowFlagsArray);
/* RayTracer.cu:2927 ShadowRayColoringWrapper */
rsh_status rsh_shadow_ray_coloring(rsh_ctx* ctx, void* stream, unsigned int windowWidth, unsigned int windowHeight,
                                   unsigned int lightTotal, rsh_float3 cameraPosition, unsigned int* shadowFlagsArray,
                                   unsigned int* pixelBufferObject);
/* RayTracer.cu:2959 ReflectionRayPreparationWrapper */
rsh_status rsh_reflection_ray_preparation(rsh_ctx* ctx, void* stream, unsigned int windowWidth,
                                          unsigned int windowHeight, unsigned int* intersectionTimeArray);
/* RayTracer.cu:2977 ReflectionRayIntersectionWrapper */
rsh_status rsh_reflection_ray_intersection(rsh_ctx* ctx, void* stream, rsh_float3* rayArray,
                                           unsigned int* sortedRayIndexKeysArray,
                                           unsigned int* sortedRayIndexValuesArray, unsigned int* hierarchyHitsArray,
                                           rsh_float4* trianglePositionsArray, unsigned int hitTotal,
                                           unsigned int rayTotal, unsigned int triangleOffset,
                                           unsigned int windowWidth, unsigned int windowHeight,
                                           unsigned int lightTotal, rsh_float3 cameraPosition,
                                           unsigned int* intersectionTimeArray);
/* RayTracer.cu:3033 ReflectionRayColoringWrapper */
rsh_status rsh_reflection_ray_coloring(rsh_ctx* ctx, void* stream, rsh_float4* trianglePositionsArray,
                                       rsh_float4* triangleNormalsArray, unsigned int windowWidth,
                                       unsigned int windowHeight, unsigned int lightTotal,
                                       rsh_float3 cameraPosition, bool createRays,
                                       unsigned int* intersectionTimeArray, rsh_float3* rayArray,
                                       unsigned int* headFlagsArray, unsigned int* rayIndexKeysArray,
                                       unsigned int* rayIndexValuesArray, unsigned int* pixelBufferObject);

#ifdef __cplusplus
}
#endif

#endif /* RSH_CTX_H_ */
