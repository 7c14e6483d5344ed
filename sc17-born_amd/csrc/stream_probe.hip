// stream_probe: the HBM rate that plain streaming kernels reach on this GPU,
// the practical ceiling the step kernels' traffic is compared with.
//
// Each kernel moves whole 1 GiB buffers once per launch (one element per
// thread, no reuse); the best and mean of 10 timed launches (HIP events) are
// printed as GB/s of bytes moved.  Patterns: a copy with 16-byte lanes, the
// same with non-temporal loads and stores, a copy with 4-byte lanes, a read
// and a write alone, six 4-byte input streams and two output streams (the
// shape of the FP32 forward step's pointwise traffic), and hipMemcpyAsync.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                 \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "%s at line %d\n", hipGetErrorString(e_), __LINE__);    \
      return 1;                                                               \
    }                                                                         \
  } while (0)

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void copy16(const v4f *__restrict__ a, v4f *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void copy16_nt(const v4f *__restrict__ a, v4f *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ void copy4(const float *__restrict__ a, float *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
// the sum is never equal to the flag value (all inputs are zero): the
// compiler cannot drop the loads, and nothing is written
__global__ void read16(const v4f *__restrict__ a, float *__restrict__ flag, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const v4f v = a[i];
    const float s = v.x + v.y + v.z + v.w;
    if (s == 1234.5f) flag[0] = s;
  }
}
__global__ void write16(v4f *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = v4f{1.f, 2.f, 3.f, 4.f};
}
__global__ void six_in_two_out(const float *__restrict__ a0, const float *__restrict__ a1,
                               const float *__restrict__ a2, const float *__restrict__ a3,
                               const float *__restrict__ a4, const float *__restrict__ a5,
                               float *__restrict__ o0, float *__restrict__ o1, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    o0[i] = a0[i] + a1[i] * a4[i];
    o1[i] = a2[i] + a3[i] * a5[i];
  }
}

int main() {
  const long bytes = 1L << 30;
  const long n16 = bytes / 16, n4 = bytes / 4;
  const int T = 256;
  float *buf[8];
  for (int k = 0; k < 8; k++) {
    CK(hipMalloc(&buf[k], bytes));
    CK(hipMemset(buf[k], 0, bytes));
  }
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto run = [&](const char *name, double moved, auto launch) -> int {
    launch();
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    float best = 1e30f, total = 0.f;
    for (int r = 0; r < 10; r++) {
      CK(hipEventRecord(e0, 0));
      launch();
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      best = ms < best ? ms : best;
      total += ms;
    }
    printf("%-16s %6.3f ms best  %5.0f GB/s best  %5.0f GB/s mean\n", name, best,
           moved / best / 1e6, moved / (total / 10) / 1e6);
    fflush(stdout);
    return 0;
  };
  const v4f *a16 = (const v4f *)buf[0];
  v4f *b16 = (v4f *)buf[1];
  const unsigned g16 = (unsigned)((n16 + T - 1) / T), g4 = (unsigned)((n4 + T - 1) / T);
  if (run("copy 16 B/lane", 2.0 * bytes, [&] { copy16<<<g16, T>>>(a16, b16, n16); })) return 1;
  if (run("copy 16 B/lane nt", 2.0 * bytes, [&] { copy16_nt<<<g16, T>>>(a16, b16, n16); }))
    return 1;
  if (run("copy 4 B/lane", 2.0 * bytes, [&] { copy4<<<g4, T>>>(buf[0], buf[1], n4); })) return 1;
  if (run("read 16 B/lane", 1.0 * bytes, [&] { read16<<<g16, T>>>(a16, buf[2], n16); }))
    return 1;
  if (run("write 16 B/lane", 1.0 * bytes, [&] { write16<<<g16, T>>>(b16, n16); })) return 1;
  if (run("6 in 2 out 4 B", 8.0 * bytes, [&] {
        six_in_two_out<<<g4, T>>>(buf[0], buf[1], buf[2], buf[3], buf[4], buf[5], buf[6],
                                  buf[7], n4);
      }))
    return 1;
  if (run("hipMemcpyAsync", 2.0 * bytes,
          [&] { (void)hipMemcpyAsync(buf[1], buf[0], bytes, hipMemcpyDeviceToDevice, 0); }))
    return 1;
  for (int k = 0; k < 8; k++) CK(hipFree(buf[k]));
  printf("stream probe done\n");
  return 0;
}
