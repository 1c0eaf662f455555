// lgpu_noalloc.hip -- CalcTimeConstraintsForElems (lulesh.cc:2447-2596) on
// device arrays the caller holds, with no device allocation by the library.
//
// The Courant/hydro minima need one pair of block partials per workgroup.
// They live in a file-scope __device__ array of fixed size (kNaMaxBlocks
// pairs, 32 KiB), which every GPU the library runs on keeps until the process
// ends.  Its device address is taken once per device with
// hipGetSymbolAddress and only that address is handed to kernels: in host
// code the array's name is a host-side stand-in, not a device pointer.
//
// Every call on one GPU shares that GPU's array.  Each call's partial kernel
// writes every slot its final kernel reads, so nothing carries over between
// calls; calls that overlap on one GPU (two streams, two threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lulesh_gpu.h"

namespace {

constexpr int kNaBlock = 256;       // threads per workgroup
constexpr int kNaMaxBlocks = 2048;  // workgroups of the partial kernel (grid-stride beyond)
constexpr int kNaMaxDevices = 64;

// Block partials: [2*b] Courant, [2*b+1] hydro.
__device__ __align__(256) double scratch[2 * kNaMaxBlocks];

double *g_scratch[kNaMaxDevices] = {};
std::mutex g_scratchMutex;

// Device address of this device's scratch array; nullptr on failure.
double *device_scratch(int dev)
{
   std::lock_guard<std::mutex> lock(g_scratchMutex);
   if (!g_scratch[dev]) {
      void *p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
         (void)hipGetLastError();
         return nullptr;
      }
      g_scratch[dev] = static_cast<double *>(p);
   }
   return g_scratch[dev];
}

__device__ __forceinline__ double na_wave_min(double v)
{
#pragma unroll
   for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      v = (o < v) ? o : v;
   }
   return v;
}

// Block minimum of (a, b), valid in thread 0.
__device__ __forceinline__ void na_block_min2(double &a, double &b)
{
   __shared__ double sa[kNaBlock / 64], sb[kNaBlock / 64];
   a = na_wave_min(a);
   b = na_wave_min(b);
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   if (lane == 0) {
      sa[w] = a;
      sb[w] = b;
   }
   __syncthreads();
   if (threadIdx.x == 0)
      for (int i = 1; i < kNaBlock / 64; ++i) {
         a = (sa[i] < a) ? sa[i] : a;
         b = (sb[i] < b) ? sb[i] : b;
      }
}

// Courant (lulesh.cc:2447-2512) and hydro (lulesh.cc:2517-2572) candidates
// of the listed elements, min-reduced per workgroup into part[2*blockIdx].
// Candidates are taken only where vdov != 0, as in the reference; the
// minimum does not depend on the order of the elements.
__global__ void __launch_bounds__(kNaBlock)
k_tc_partial(const double *__restrict__ ss, const double *__restrict__ vdov,
             const double *__restrict__ arealg, const int32_t *__restrict__ list, int32_t n,
             double qqc2, double dvovmax, double *__restrict__ part)
{
   double dtc = 1.0e+20, dth = 1.0e+20;
   const int64_t stride = (int64_t)gridDim.x * kNaBlock;
   for (int64_t i = (int64_t)blockIdx.x * kNaBlock + threadIdx.x; i < n; i += stride) {
      const int32_t e = list ? list[i] : (int32_t)i;
      const double vd = vdov[e];
      if (vd == 0.) continue;
      const double ar = arealg[e];
      const double sse = ss[e];
      double dtf = sse * sse;
      if (vd < 0.) dtf = dtf + qqc2 * ar * ar * vd * vd;
      dtf = sqrt(dtf);
      dtf = ar / dtf;
      if (dtf < dtc) dtc = dtf;
      const double dtdvov = dvovmax / (fabs(vd) + 1.e-20);
      if (dtdvov < dth) dth = dtdvov;
   }
   na_block_min2(dtc, dth);
   if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = dtc;
      part[2 * blockIdx.x + 1] = dth;
   }
}

// Minimum over the nPart partial pairs into out[0..1].
__global__ void __launch_bounds__(kNaBlock)
k_tc_final(const double *__restrict__ part, int nPart, double *__restrict__ out)
{
   double c = 1.0e+20, h = 1.0e+20;
   for (int b = threadIdx.x; b < nPart; b += kNaBlock) {
      const double pc = part[2 * b], ph = part[2 * b + 1];
      c = (pc < c) ? pc : c;
      h = (ph < h) ? ph : h;
   }
   na_block_min2(c, h);
   if (threadIdx.x == 0) {
      out[0] = c;
      out[1] = h;
   }
}

} // namespace

extern "C" int lgpu_time_constraints_noalloc(const double *ss, const double *vdov,
                                             const double *arealg, const int32_t *elemlist,
                                             int32_t numElem, double qqc, double dvovmax,
                                             double *dtOut, void *stream)
{
   if (!ss || !vdov || !arealg || !dtOut || numElem < 0) return LGPU_EINVAL;
   int dev = 0;
   if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) {
      (void)hipGetLastError();
      return LGPU_EDEVICE;
   }
   double *part = device_scratch(dev);
   if (!part) return LGPU_EDEVICE;
   const hipStream_t st = static_cast<hipStream_t>(stream);
   const int64_t want = ((int64_t)numElem + kNaBlock - 1) / kNaBlock;
   const int grid = (int)(want < kNaMaxBlocks ? want : kNaMaxBlocks);
   const double qqc2 = 64.0 * qqc * qqc;
   if (grid > 0)
      hipLaunchKernelGGL(k_tc_partial, dim3(grid), dim3(kNaBlock), 0, st, ss, vdov, arealg,
                         elemlist, numElem, qqc2, dvovmax, part);
   hipLaunchKernelGGL(k_tc_final, dim3(1), dim3(kNaBlock), 0, st, part, grid, dtOut);
   if (hipGetLastError() != hipSuccess) return LGPU_EDEVICE;
   return LGPU_OK;
}
