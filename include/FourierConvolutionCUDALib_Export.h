/* FourierConvolutionCUDALib_Export.h -- symbol visibility for the MI355X
 * build of libFourierConvolutionCUDALib.
 *
 * Replaces src/FourierConvolutionCUDALib_Export.h (reference), which CMake's
 * GenerateExportHeader writes for MSVC (src/CMakeLists.txt:17-25).  On Linux
 * (the only platform this library targets) exported symbols need no
 * decoration beyond default visibility, so the same macro names are kept
 * and expand to a visibility attribute; FourierConvolutionCUDALib_BUILT_AS_STATIC
 * keeps its meaning for the static archive.
 */
#ifndef FourierConvolutionCUDALib_EXPORT_H
#define FourierConvolutionCUDALib_EXPORT_H

#ifdef FourierConvolutionCUDALib_BUILT_AS_STATIC
#define FourierConvolutionCUDALib_EXPORT
#define FOURIERCONVOLUTIONCUDALIB_NO_EXPORT
#else
#ifndef FourierConvolutionCUDALib_EXPORT
#define FourierConvolutionCUDALib_EXPORT __attribute__((visibility("default")))
#endif
#ifndef FOURIERCONVOLUTIONCUDALIB_NO_EXPORT
#define FOURIERCONVOLUTIONCUDALIB_NO_EXPORT __attribute__((visibility("hidden")))
#endif
#endif

#ifndef FOURIERCONVOLUTIONCUDALIB_DEPRECATED
#define FOURIERCONVOLUTIONCUDALIB_DEPRECATED __attribute__((deprecated))
#endif

#endif /* FourierConvolutionCUDALib_EXPORT_H */
