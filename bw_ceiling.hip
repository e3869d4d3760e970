// Memory-bandwidth ceilings of this GPU for the access shapes of the fused
// lattice pass, used as the reference line next to its roofline figure.
//
// Kernels (each timed with hipEvents over 20 launches after 3 warm-ups, at
// several working-set sizes; GB/s = bytes read + written / time):
//   copy4        out[i] = in[i], 16 bytes per lane (the plain-copy ceiling)
//   read4        16-byte loads only
//   soa19        19 planes read, 19 planes written (out of place), 4 bytes/lane
//   soa19_inpl   19 planes read then written back in place (the AA even pass)
//   soa19_shift  in place, loads from c - s_i and stores to c + s_i with the
//                D3Q19 offsets of a 128 x 32 plane (the AA odd pass shape)
//   soa19_inpl4  soa19_inpl with 4 consecutive cells per thread: one 16-byte
//                load and store per plane
//   soa19_shift4 soa19_shift with 4 consecutive cells per thread: 16-byte
//                accesses at the D3Q19 offsets (4-byte aligned only)
// Prints one JSON line per kernel and size.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                        \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      exit(1);                                                                       \
    }                                                                                \
  } while (0)

__global__ void copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

__global__ void read4(const float4* __restrict__ a, float* __restrict__ out, size_t n4) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const float4 v = i < n4 ? a[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float s = v.x + v.y + v.z + v.w;
  if (s == -1.0f) out[0] = s;  // the buffer holds zeros: never taken, keeps the loads
}

__global__ void soa19(const float* __restrict__ a, float* __restrict__ b, uint32_t N) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float v[19];
#pragma unroll
  for (int p = 0; p < 19; ++p) v[p] = a[(size_t)p * N + c];
#pragma unroll
  for (int p = 0; p < 19; ++p) b[(size_t)p * N + c] = v[p] * 0.5f;
}

__global__ void soa19_inpl(float* __restrict__ a, uint32_t N) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float v[19];
#pragma unroll
  for (int p = 0; p < 19; ++p) v[p] = a[(size_t)p * N + c];
#pragma unroll
  for (int p = 0; p < 19; ++p) a[(size_t)(18 - p) * N + c] = v[p] * 0.5f;
}

// D3Q19 offsets (rest, faces, edges) on a lattice with ny = 128, nz = 32.
__constant__ int kShift[19] = {0,    4096, -4096, 32,    -32,   1,  -1,  4128, -4128, 4064,
                               -4064, 4097, -4097, 4095, -4095, 33, -33, 31,   -31};
constexpr uint32_t kMaxShift = 4128;

__global__ void soa19_shift(float* __restrict__ a, uint32_t N, uint32_t lo, uint32_t hi) {
  const uint32_t c = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= hi) return;  // lo, hi keep c +- kMaxShift inside [0, N)
  float v[19];
#pragma unroll
  for (int p = 0; p < 19; ++p) {
    const int q = p == 0 ? 0 : ((p & 1) ? p + 1 : p - 1);
    v[p] = a[(size_t)q * N + (c - kShift[p])];
  }
#pragma unroll
  for (int p = 0; p < 19; ++p) a[(size_t)p * N + (c + kShift[p])] = v[p] * 0.5f;
}

// 16 bytes with 4-byte alignment: the shifted accesses start at any cell.
typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));

// N4: plane stride in floats, a multiple of 4; thread t owns cells 4t .. 4t+3.
__global__ void soa19_inpl4(float* __restrict__ a, uint32_t N4) {
  const uint32_t c = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= N4) return;
  f4a4 v[19];
#pragma unroll
  for (int p = 0; p < 19; ++p) v[p] = *(const f4a4*)(a + (size_t)p * N4 + c);
#pragma unroll
  for (int p = 0; p < 19; ++p) *(f4a4*)(a + (size_t)(18 - p) * N4 + c) = v[p] * 0.5f;
}

__global__ void soa19_shift4(float* __restrict__ a, uint32_t N4, uint32_t lo, uint32_t hi) {
  const uint32_t c = lo + 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= hi) return;  // lo, hi keep c - kMaxShift and c + 3 + kMaxShift inside [0, N4)
  f4a4 v[19];
#pragma unroll
  for (int p = 0; p < 19; ++p) {
    const int q = p == 0 ? 0 : ((p & 1) ? p + 1 : p - 1);
    v[p] = *(const f4a4*)(a + (size_t)q * N4 + (c - kShift[p]));
  }
#pragma unroll
  for (int p = 0; p < 19; ++p) *(f4a4*)(a + (size_t)p * N4 + (c + kShift[p])) = v[p] * 0.5f;
}

int main() {
  const size_t sizes_mb[] = {92, 512, 2048};
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  for (size_t mb : sizes_mb) {
    const size_t n = mb * (size_t(1) << 20) / 4;  // floats per buffer
    float *a, *b;
    CK(hipMalloc(&a, n * 4));
    CK(hipMalloc(&b, n * 4));
    CK(hipMemset(a, 0, n * 4));
    CK(hipMemset(b, 0, n * 4));
    const int reps = 20;
    auto run = [&](const char* name, double bytes, auto launch) {
      for (int r = 0; r < 3; ++r) launch();
      CK(hipEventRecord(e0));
      for (int r = 0; r < reps; ++r) launch();
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      CK(hipGetLastError());
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      printf("{\"mode\": \"ceiling\", \"kernel\": \"%s\", \"MiB_per_buffer\": %zu, \"GBs\": %.1f}\n", name, mb,
             bytes / (ms / reps * 1e-3) / 1e9);
      fflush(stdout);
    };
    const size_t n4 = n / 4;
    const unsigned g4 = (unsigned)((n4 + 255) / 256);
    run("copy4", 8.0 * n, [&] { copy4<<<g4, 256>>>((const float4*)a, (float4*)b, n4); });
    run("read4", 4.0 * n, [&] { read4<<<g4, 256>>>((const float4*)a, b, n4); });
    const uint32_t N = (uint32_t)(n / 19);
    const unsigned gs = (N + 255) / 256;
    run("soa19", 8.0 * 19 * N, [&] { soa19<<<gs, 256>>>(a, b, N); });
    run("soa19_inpl", 8.0 * 19 * N, [&] { soa19_inpl<<<gs, 256>>>(a, N); });
    const uint32_t lo = 2 * kMaxShift, hi = N - 2 * kMaxShift;
    run("soa19_shift", 8.0 * 19 * (hi - lo), [&] { soa19_shift<<<(hi - lo + 255) / 256, 256>>>(a, N, lo, hi); });
    const uint32_t N4 = N & ~3u;  // 19 * N4 <= n
    const unsigned g4c = (N4 / 4 + 255) / 256;
    run("soa19_inpl4", 8.0 * 19 * N4, [&] { soa19_inpl4<<<g4c, 256>>>(a, N4); });
    const uint32_t lo4 = 2 * kMaxShift, hi4 = N4 - 2 * kMaxShift;  // both multiples of 4
    run("soa19_shift4", 8.0 * 19 * (hi4 - lo4),
        [&] { soa19_shift4<<<((hi4 - lo4) / 4 + 255) / 256, 256>>>(a, N4, lo4, hi4); });
    CK(hipFree(a));
    CK(hipFree(b));
  }
  return 0;
}
