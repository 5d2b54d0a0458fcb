/* consumer_io.h -- helpers shared by the consumer programs. */
#ifndef CONSUMER_IO_H
#define CONSUMER_IO_H
#ifdef __cplusplus
extern "C" {
#endif
float* read_floats(const char* path, long n);
int write_floats(const char* path, const float* p, long n);
int header_constants_ok(void);
#ifdef __cplusplus
}
#endif
#endif
