/* The reference's outer C entry point (app/System_dependent/Linux/c_wrapper.h:
 * 10-15), served by the MI355X library (libmobilert.so).
 *
 * shader: 1 Whitted, 2 PathTracer, 3 DepthMap (maxPoint (1, 1, 1), as for the
 *         reference's unit scenes), 4 DiffuseMaterial, anything else NoShadows
 *         (c_wrapper.cpp:444-486).
 * scene:  0 the Cornell box of this build's generator (34 triangles, 2 area
 *         lights, camera at (0, 0, -3.4)), 1 the 1M-triangle soup (990k
 *         diffuse, 10k area lights), any other id with pathObj set the OBJ /
 *         MTL files (Components::OBJLoader) with the reference's default
 *         camera or its CornellBox pose; the reference's other demo scenes are
 *         refused.
 * accelerator: 0 Naive, 1 RegularGrid, 2 BVH.
 * threads is accepted and ignored; async renders on a detached host thread
 * and returns at once, as the reference does (the bitmap must outlive the
 * render).  With printStdOut the timings and scene counts are printed.
 * Failures print a message to stderr and leave the bitmap untouched (the
 * reference exits instead). */
#pragma once

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C"
#endif
void RayTrace(uint32_t* bitmap, int32_t width, int32_t height, int32_t threads, int32_t shader, int32_t scene,
              int32_t samplesPixel, int32_t samplesLight, int32_t repeats, int32_t accelerator, bool printStdOut,
              bool async, const char* pathObj, const char* pathMtl);
