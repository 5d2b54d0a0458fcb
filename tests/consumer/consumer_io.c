/* consumer_io.c -- raw float32 file I/O for the consumer programs.  A second
 * C translation unit that includes the public header, so the header's
 * file-scope constants are compiled into more than one object. */
#include <stdio.h>
#include <stdlib.h>

#include "consumer_io.h"
#include "convolution3Dfft.h"

float* read_floats(const char* path, long n) {
  FILE* f = fopen(path, "rb");
  float* buf;
  if (!f) return NULL;
  buf = (float*)malloc(sizeof(float) * (size_t)n);
  if (buf && fread(buf, sizeof(float), (size_t)n, f) != (size_t)n) {
    free(buf);
    buf = NULL;
  }
  fclose(f);
  return buf;
}

int write_floats(const char* path, const float* p, long n) {
  FILE* f = fopen(path, "wb");
  size_t w;
  if (!f) return 1;
  w = fwrite(p, sizeof(float), (size_t)n, f);
  fclose(f);
  return w == (size_t)n ? 0 : 1;
}

int header_constants_ok(void) {
  /* src/convolution3Dfft.h:17-21 */
  return MAX_THREADS_CUDA == 1024 && MAX_BLOCKS_CUDA == 65535 && dimsImage == 3 &&
         sizeof(imageType) == 4;
}
