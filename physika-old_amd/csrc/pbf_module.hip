// Module-granular kernels: the reference modules' passes over a caller CSR
// list, in original particle order (ParticleIntegrator, SummationDensity,
// DensityPBD with its pairwise atomic scatter).
#include "pbf_device.h"

namespace pbf {

// ParticleIntegrator::integrate: K_UpdateVelocity then K_UpdatePosition
// (ParticleIntegrator.cu:60-70,103-112), each multiply and add rounded
// separately as the source writes them.
__global__ __launch_bounds__(kBlock) void k_integrate(float* __restrict__ x, float* __restrict__ v,
                                                      const float* __restrict__ f, int n, PhysConst pc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = 3 * (size_t)i;
  float vx, vy, vz;
  if (f) {
    vx = v[b] + (f[b] + pc.g[0]) * pc.dt;
    vy = v[b + 1] + (f[b + 1] + pc.g[1]) * pc.dt;
    vz = v[b + 2] + (f[b + 2] + pc.g[2]) * pc.dt;
  } else {
    vx = v[b] + pc.gdt[0];
    vy = v[b + 1] + pc.gdt[1];
    vz = v[b + 2] + pc.gdt[2];
  }
  v[b] = vx;
  v[b + 1] = vy;
  v[b + 2] = vz;
  x[b] = x[b] + vx * pc.dt;
  x[b + 1] = x[b + 1] + vy * pc.dt;
  x[b + 2] = x[b + 2] + vz * pc.dt;
}

void launch_integrate(float* x, float* v, const float* f, int n, const PhysConst& pc, hipStream_t st) {
  k_integrate<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, v, f, n, pc);
}

// K_ComputeDensity (SummationDensity.cu:13-36) over a CSR list.
__global__ __launch_bounds__(kBlock) void k_csr_density(const float* __restrict__ x, Csr nb,
                                                        float* __restrict__ rho, PhysConst pc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb.n) return;
  const float4 xi = load3(x, i);
  float acc = 0.f;
  for (int k = nb.begin(i), e = nb.end(i); k < e; ++k) {
    const PairGeom p = pair_geom(xi, load3(x, nb.elems[k]), pc.inv_h);
    acc += p.t * p.t * p.t;
  }
  rho[i] = acc * pc.cW;
}

void launch_csr_density(const float* x, const Csr& nb, float* rho, const PhysConst& pc, hipStream_t st) {
  k_csr_density<<<(nb.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, nb, rho, pc);
}

__global__ __launch_bounds__(kBlock) void k_aos_to_f4(const float* __restrict__ x, float4* __restrict__ xa, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xa[i] = load3(x, i);
}

void launch_aos_to_f4(const float* x, float4* xa, int n, hipStream_t st) {
  k_aos_to_f4<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, xa, n);
}

// The copy, and the displacement sums cleared (m_deltaPos.reset(),
// DensityPBD.cu:315) for scratch that the previous call left as it was.
__global__ __launch_bounds__(kBlock) void k_aos_to_f4_clear(const float* __restrict__ x, float4* __restrict__ xa,
                                                            float* __restrict__ dpos, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xa[i] = load3(x, i);
  const size_t b = 3 * (size_t)i;
  dpos[b] = 0.f;
  dpos[b + 1] = 0.f;
  dpos[b + 2] = 0.f;
}

void launch_aos_to_f4_clear(const float* x, float4* xa, float* dpos, int n, hipStream_t st) {
  k_aos_to_f4_clear<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, xa, dpos, n);
}

// Density and K_ComputeLambdas (DensityPBD.cu:46-90) over a CSR list; lambda
// goes to xa[i].w. With minv (inverse masses) the massInv variant
// (DensityPBD.cu:93-131): S = sum_j |g|^2 m_j^-1 + |sum g|^2 m_i^-1.
__global__ __launch_bounds__(kBlock) void k_csr_lambda(float4* __restrict__ xa, Csr nb, float* __restrict__ rho,
                                                       PhysConst pc, const float* __restrict__ minv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb.n) return;
  const float4 xi = xa[i];
  float acc = 0.f, sgx = 0.f, sgy = 0.f, sgz = 0.f, sgg = 0.f;
  for (int k = nb.begin(i), e = nb.end(i); k < e; ++k) {
    const int j = nb.elems[k];
    const PairGeom p = pair_geom(xi, xa[j], pc.inv_h);
    const float t2 = p.t * p.t;
    acc += t2 * p.t;
    const float gq = p.r > pc.pair_eps ? t2 * p.rinv : 0.0f;
    const float gx = gq * p.dx, gy = gq * p.dy, gz = gq * p.dz;
    sgx += gx;
    sgy += gy;
    sgz += gz;
    const float gg = gx * gx + gy * gy + gz * gz;
    sgg += minv ? gg * minv[j] : gg;
  }
  acc *= pc.cW;
  const float G2 = sgx * sgx + sgy * sgy + sgz * sgz;
  const float S = (pc.cG * pc.cG) * (sgg + (minv ? G2 * minv[i] : G2));
  float lam = -(acc - pc.lam_rest) / (S + pc.lam_eps);
  lam = lam > 0.0f ? 0.0f : lam;
  reinterpret_cast<float*>(xa + i)[3] = lam;
  if (rho) rho[i] = acc;
}

void launch_csr_lambda(float4* xa, const Csr& nb, float* rho, const PhysConst& pc, hipStream_t st,
                       const float* minv) {
  k_csr_lambda<<<(nb.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, nb, rho, pc, minv);
}

// K_ComputeDisplacement (DensityPBD.cu:135-179): pairwise scatter with float
// atomics, valid for any list (symmetric or not), order nondeterministic as in
// the reference. With minv the massInv variant (DensityPBD.cu:182-224):
// dp_ij m_i^-1 to i, -dp_ij m_j^-1 to j.
__global__ __launch_bounds__(kBlock) void k_csr_dx_atomic(const float4* __restrict__ xa, Csr nb,
                                                          float* __restrict__ dpos, PhysConst pc,
                                                          const float* __restrict__ minv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb.n) return;
  const float4 xi = xa[i];
  const float scale = pc.corr * pc.cG;
  const float mi = minv ? minv[i] : 1.0f;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int k = nb.begin(i), e = nb.end(i); k < e; ++k) {
    const int j = nb.elems[k];
    const float4 xj = xa[j];
    const PairGeom p = pair_geom(xi, xj, pc.inv_h);
    if (!(p.r > pc.pair_eps)) continue;
    const float c = scale * (xi.w + xj.w) * (p.t * p.t) * p.rinv;
    const float px = c * p.dx, py = c * p.dy, pz = c * p.dz;
    const float mj = minv ? minv[j] : 1.0f;
    ax += px * mi;
    ay += py * mi;
    az += pz * mi;
    atomicAdd(&dpos[3 * (size_t)j], -px * mj);
    atomicAdd(&dpos[3 * (size_t)j + 1], -py * mj);
    atomicAdd(&dpos[3 * (size_t)j + 2], -pz * mj);
  }
  atomicAdd(&dpos[3 * (size_t)i], ax);
  atomicAdd(&dpos[3 * (size_t)i + 1], ay);
  atomicAdd(&dpos[3 * (size_t)i + 2], az);
}

void launch_csr_dx_atomic(const float4* xa, const Csr& nb, float* dpos, const PhysConst& pc, hipStream_t st,
                          const float* minv) {
  k_csr_dx_atomic<<<(nb.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, nb, dpos, pc, minv);
}

// K_UpdatePosition (DensityPBD.cu:227-237); clears dpos for the next round
// (m_deltaPos.reset(), DensityPBD.cu:315).
__global__ __launch_bounds__(kBlock) void k_csr_apply(float4* __restrict__ xa, float* __restrict__ dpos, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = 3 * (size_t)i;
  float4 p = xa[i];
  p.x += dpos[b];
  p.y += dpos[b + 1];
  p.z += dpos[b + 2];
  xa[i] = p;
  dpos[b] = 0.f;
  dpos[b + 1] = 0.f;
  dpos[b + 2] = 0.f;
}

void launch_csr_apply(float4* xa, float* dpos, int n, hipStream_t st) {
  k_csr_apply<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, dpos, n);
}

// DP_UpdateVelocity (DensityPBD.cu:368-378) with x_old = the caller's x, then
// x := the constrained positions.
__global__ __launch_bounds__(kBlock) void k_csr_finish(const float4* __restrict__ xa, float* __restrict__ x,
                                                       float* __restrict__ v, int n, float dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = 3 * (size_t)i;
  const float4 p = xa[i];
  v[b] += (p.x - x[b]) / dt;
  v[b + 1] += (p.y - x[b + 1]) / dt;
  v[b + 2] += (p.z - x[b + 2]) / dt;
  x[b] = p.x;
  x[b + 1] = p.y;
  x[b + 2] = p.z;
}

void launch_csr_finish(const float4* xa, float* x, float* v, int n, float dt, hipStream_t st) {
  k_csr_finish<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, x, v, n, dt);
}

}  // namespace pbf
