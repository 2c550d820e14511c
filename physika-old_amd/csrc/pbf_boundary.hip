// The two modules around the PBD solve in a PhysIKA fluid scene:
//  - ImplicitViscosity (Dynamics/ParticleSystem/ImplicitViscosity.cu:9-128),
//    the last constraint of PositionBasedFluidModel::step
//    (PositionBasedFluidModel.cpp:69-75,98): Jacobi iterations over the step's
//    neighbor lists, in sorted order on the fused path and over a caller CSR
//    list on the module path;
//  - the box boundary of StaticBoundary::loadCube (StaticBoundary.cpp:72-77):
//    a sampled signed distance field of the box (DistanceField3D.cu:90-129,
//    194-201,283-291) and K_ConstrainSDF (BoundaryConstraint.cu:30-67), run on
//    the particles before each fluid step.
#include "pbf_nbr.h"

namespace pbf {

// ---------------------------------------------------------------------------
// Box field. CUDA's float min returns the non-NaN operand, as fminf does.

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
#pragma clang fp contract(off)
  const float tx = ax * bx, ty = ay * by, tz = az * bz;
  return (tx + ty) + tz;
}

// DistanceToSegment (DistanceField3D.cu:14-32).
__device__ float seg_distance(float3 p, float3 a, float3 b) {
#pragma clang fp contract(off)
  float sx = b.x - a.x, sy = b.y - a.y, sz = b.z - a.z;
  const float e1x = p.x - a.x, e1y = p.y - a.y, e1z = p.z - a.z;
  const float e2x = p.x - b.x, e2y = p.y - b.y, e2z = p.z - b.z;
  if (dot3(e1x, e1y, e1z, sx, sy, sz) < 0.0f) return sqrtf(dot3(e1x, e1y, e1z, e1x, e1y, e1z));
  if (dot3(e2x, e2y, e2z, -sx, -sy, -sz) < 0.0f) return sqrtf(dot3(e2x, e2y, e2z, e2x, e2y, e2z));
  const float length1 = dot3(e1x, e1y, e1z, e1x, e1y, e1z);
  // Vector::normalize: x * (1/sqrt(dot(x,x))) when |x| > FLT_EPSILON, else 0
  const float ss = dot3(sx, sy, sz, sx, sy, sz);
  if (sqrtf(ss) > 1.1920928955078125e-07f) {
    const float is = 1.0f / sqrtf(ss);
    sx *= is;
    sy *= is;
    sz *= is;
  } else {
    sx = sy = sz = 0.0f;
  }
  const float length2 = dot3(e1x, e1y, e1z, sx, sy, sz);
  return sqrtf(length1 - length2 * length2);
}

// DistanceToSqure (DistanceField3D.cu:34-88): distance to the face rectangle
// with corner lo and opposite corner hi, normal along `axis`.
__device__ float square_distance(float3 p, float3 lo, float3 hi, int axis) {
#pragma clang fp contract(off)
  float3 c1, c2, c3, c4;
  float l0, l1, h0, h1, q0, q1, dn;
  if (axis == 0) {
    c1 = make_float3(lo.x, lo.y, lo.z);
    c2 = make_float3(lo.x, hi.y, lo.z);
    c3 = make_float3(lo.x, hi.y, hi.z);
    c4 = make_float3(lo.x, lo.y, hi.z);
    l0 = lo.y; l1 = lo.z; h0 = hi.y; h1 = hi.z; q0 = p.y; q1 = p.z;
    dn = dot3(1.0f, 0.0f, 0.0f, p.x - c1.x, p.y - c1.y, p.z - c1.z);
  } else if (axis == 1) {
    c1 = make_float3(lo.x, lo.y, lo.z);
    c2 = make_float3(lo.x, lo.y, hi.z);
    c3 = make_float3(hi.x, lo.y, hi.z);
    c4 = make_float3(hi.x, lo.y, lo.z);
    l0 = lo.x; l1 = lo.z; h0 = hi.x; h1 = hi.z; q0 = p.x; q1 = p.z;
    dn = dot3(0.0f, 1.0f, 0.0f, p.x - c1.x, p.y - c1.y, p.z - c1.z);
  } else {
    c1 = make_float3(lo.x, lo.y, lo.z);
    c2 = make_float3(hi.x, lo.y, lo.z);
    c3 = make_float3(hi.x, hi.y, lo.z);
    c4 = make_float3(lo.x, hi.y, lo.z);
    l0 = lo.x; l1 = lo.y; h0 = hi.x; h1 = hi.y; q0 = p.x; q1 = p.y;
    dn = dot3(0.0f, 0.0f, 1.0f, p.x - c1.x, p.y - c1.y, p.z - c1.z);
  }
  const float d1 = seg_distance(p, c1, c2), d2 = seg_distance(p, c2, c3);
  const float d3 = seg_distance(p, c3, c4), d4 = seg_distance(p, c4, c1);
  if (q0 < h0 && q0 > l0 && q1 < h1 && q1 > l1) return fabsf(dn);
  return fminf(fminf(d1, d2), fminf(d3, d4));
}

// DistanceToBox (DistanceField3D.cu:90-129).
__device__ float box_distance(float3 p, float3 lo, float3 hi) {
  if (p.x < hi.x && p.x > lo.x && p.y < hi.y && p.y > lo.y && p.z < hi.z && p.z > lo.z) {
    const float dx = fminf(fabsf(p.x - hi.x), fabsf(p.x - lo.x));
    const float dy = fminf(fabsf(p.y - hi.y), fabsf(p.y - lo.y));
    const float dz = fminf(fabsf(p.z - hi.z), fabsf(p.z - lo.z));
    return fminf(fminf(dx, dy), dz);
  }
  const float3 c0 = lo, c1 = make_float3(hi.x, lo.y, lo.z), c2 = make_float3(hi.x, hi.y, lo.z);
  const float3 c3 = make_float3(lo.x, hi.y, lo.z), c4 = make_float3(lo.x, lo.y, hi.z);
  const float3 c5 = make_float3(hi.x, lo.y, hi.z), c6 = hi, c7 = make_float3(lo.x, hi.y, hi.z);
  const float x1 = square_distance(p, c0, c7, 0), x2 = square_distance(p, c1, c6, 0);
  const float y1 = square_distance(p, c0, c5, 1), y2 = square_distance(p, c3, c6, 1);
  const float z1 = square_distance(p, c0, c2, 2), z2 = square_distance(p, c4, c6, 2);
  return -fminf(fminf(fminf(x1, x2), fminf(y1, y2)), fminf(z1, z2));
}

// K_DistanceFieldToBox (DistanceField3D.cu:268-281): sample (i,j,k) at
// left + (i,j,k)*h, x fastest; positive inside the box (inverted field).
__global__ __launch_bounds__(kBlock) void k_box_sdf(BoxField f, float3 lo, float3 hi, float sign) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)f.nx * f.ny * f.nz;
  if (t >= total) return;
  const int i = (int)(t % f.nx), j = (int)((t / f.nx) % f.ny), k = (int)(t / ((long long)f.nx * f.ny));
  const float3 p = make_float3(f.left[0] + (float)i * f.h[0], f.left[1] + (float)j * f.h[1],
                               f.left[2] + (float)k * f.h[2]);
  f.d[t] = sign * box_distance(p, lo, hi);
}

void launch_box_sdf(const BoxField& f, const float lo[3], const float hi[3], hipStream_t st) {
  const long long total = (long long)f.nx * f.ny * f.nz;
  k_box_sdf<<<(unsigned)((total + kBlock - 1) / kBlock), kBlock, 0, st>>>(
      f, make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]), 1.0f);
}

// DistanceField3D::getDistance (DistanceField3D.h:112-170) of the inverted
// field: outside the sampled space d = -100000 and no normal.
__device__ void field_distance(const BoxField& f, float px, float py, float pz, float& d, float3& nrm) {
#pragma clang fp contract(off)
  const float fx = (px - f.left[0]) * f.inv_h[0];
  const float fy = (py - f.left[1]) * f.inv_h[1];
  const float fz = (pz - f.left[2]) * f.inv_h[2];
  const float ffx = floorf(fx), ffy = floorf(fy), ffz = floorf(fz);
  if (!(ffx >= 0.0f && ffx < (float)(f.nx - 1) && ffy >= 0.0f && ffy < (float)(f.ny - 1) && ffz >= 0.0f &&
        ffz < (float)(f.nz - 1))) {
    d = -100000.0f;
    nrm = make_float3(0.0f, 0.0f, 0.0f);
    return;
  }
  const int i = (int)ffx, j = (int)ffy, k = (int)ffz;
  const float alpha = fx - (float)i, beta = fy - (float)j, gamma = fz - (float)k;
  const long long sx = 1, sy = f.nx, sz = (long long)f.nx * f.ny;
  const float* b = f.d + i * sx + j * sy + k * sz;
  const float d000 = b[0], d100 = b[sx], d010 = b[sy], d110 = b[sx + sy];
  const float d001 = b[sz], d101 = b[sx + sz], d011 = b[sy + sz], d111 = b[sx + sy + sz];
  auto lerp = [](float a, float c, float t) {
#pragma clang fp contract(off)
    return (1.0f - t) * a + t * c;
  };
  const float dx00 = lerp(d000, d100, alpha), dx10 = lerp(d010, d110, alpha), dxy0 = lerp(dx00, dx10, beta);
  const float dx01 = lerp(d001, d101, alpha), dx11 = lerp(d011, d111, alpha), dxy1 = lerp(dx01, dx11, beta);
  const float d0y0 = lerp(d000, d010, beta), d0y1 = lerp(d001, d011, beta), d0yz = lerp(d0y0, d0y1, gamma);
  const float d1y0 = lerp(d100, d110, beta), d1y1 = lerp(d101, d111, beta), d1yz = lerp(d1y0, d1y1, gamma);
  const float dx0z = lerp(dx00, dx01, gamma), dx1z = lerp(dx10, dx11, gamma);
  float nx = d0yz - d1yz, ny = dx0z - dx1z, nz = dxy0 - dxy1;
  const float nn = dot3(nx, ny, nz, nx, ny, nz);
  const float l = sqrtf(nn);
  if (l < 0.0001f) {
    nx = ny = nz = 0.0f;
  } else {
    const float is = 1.0f / sqrtf(nn);
    nx *= is;
    ny *= is;
    nz *= is;
  }
  nrm = make_float3(nx, ny, nz);
  d = (1.0f - gamma) * dxy0 + gamma * dxy1;
}

// 1e-4 times a uniform in (0, 1] per particle id: SplitMix64 of id + 1, top 24
// bits (the reference's draw is curand's, which the library does not
// reproduce).
__device__ __forceinline__ float box_jitter(int id) {
  unsigned long long z = ((unsigned long long)(unsigned)id + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return 0.0001f * ((float)((z >> 40) + 1ull) * (1.0f / 16777216.0f));
}

// K_ConstrainSDF (BoundaryConstraint.cu:30-67) on (x_in, v_in) -> (x_out,
// v_out); the particle id is its index inside its scene.
__global__ __launch_bounds__(kBlock) void k_box_constrain(BoxField f, const float* __restrict__ x_in,
                                                          const float* __restrict__ v_in, float* __restrict__ x_out,
                                                          float* __restrict__ v_out, int n,
                                                          const int* __restrict__ scene_off, int nscenes,
                                                          float keep_normal, float damp) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float px = x_in[3 * (size_t)i], py = x_in[3 * (size_t)i + 1], pz = x_in[3 * (size_t)i + 2];
  float vx = v_in[3 * (size_t)i], vy = v_in[3 * (size_t)i + 1], vz = v_in[3 * (size_t)i + 2];
  float dist;
  float3 nr;
  field_distance(f, px, py, pz, dist, nr);
  if (dist <= 0.0f) {
    int lo = 0, hi = nscenes;  // scene_off[lo] <= i < scene_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (scene_off[mid] <= i)
        lo = mid;
      else
        hi = mid;
    }
    const float olddist = -dist;
    dist = box_jitter(i - scene_off[lo]);
    const float s = olddist + dist;
    px = px - s * nr.x;
    py = py - s * nr.y;
    pz = pz - s * nr.z;
    const float vn = dot3(vx, vy, vz, nr.x, nr.y, nr.z);
    float nx = vn * nr.x, ny = vn * nr.y, nz = vn * nr.z;
    const float tx = vx - nx, ty = vy - ny, tz = vz - nz;
    if (vn > 0.0f) {
      nx = -nx;
      ny = -ny;
      nz = -nz;
    }
    nx *= keep_normal;
    ny *= keep_normal;
    nz *= keep_normal;
    vx = (nx + tx) * damp;
    vy = (ny + ty) * damp;
    vz = (nz + tz) * damp;
  }
  x_out[3 * (size_t)i] = px;
  x_out[3 * (size_t)i + 1] = py;
  x_out[3 * (size_t)i + 2] = pz;
  v_out[3 * (size_t)i] = vx;
  v_out[3 * (size_t)i + 1] = vy;
  v_out[3 * (size_t)i + 2] = vz;
}

void launch_box_constrain(const BoxField& f, const float* x_in, const float* v_in, float* x_out, float* v_out, int n,
                          const int* scene_off, int nscenes, float keep_normal, float damp, hipStream_t st) {
  k_box_constrain<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(f, x_in, v_in, x_out, v_out, n, scene_off, nscenes,
                                                               keep_normal, damp);
}

// ---------------------------------------------------------------------------
// ImplicitViscosity

struct VisPair {
  float4 x, v;
};

// Sums of one viscosity iteration for a particle (or one entry parity of it).
struct VisSums {
  float tw = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
};

// w_ij = 45/(13 pi h^3)(1 - r/h) for r > eps: adds w and w v_j.
__device__ __forceinline__ void vis_term(VisSums& a, float4 xi, const VisPair& q, const VisConst& vc) {
#pragma clang fp contract(off)
  const PairGeom g = pair_geom(xi, q.x, vc.inv_h);
  if (g.r > vc.eps) {
    const float w = vc.cw * g.t;
    a.tw += w;
    a.dx = fmaf(w, q.v.x, a.dx);
    a.dy = fmaf(w, q.v.y, a.dy);
    a.dz = fmaf(w, q.v.z, a.dz);
  }
}

// v_new = v_old/(1+b) + (sum w v_j / sum w) b/(1+b) from the total sums t; the
// last iteration writes positions and velocities back to original order.
__device__ __forceinline__ void vis_finish(int s, float4 xi, VisSums t, const float4* __restrict__ v_old,
                                           float4* __restrict__ v_next, const VisConst& vc,
                                           const int* __restrict__ perm, float* __restrict__ x_out,
                                           float* __restrict__ v_out) {
  // nothing contracted, so both kernel forms round it alike
#pragma clang fp contract(off)
  const float b = t.tw < vc.eps ? 0.0f : vc.b;
  const float tw = t.tw < vc.eps ? 1.0f : t.tw;
  const float ob = 1.0f + b;
  const float4 vo = v_old[s];
  const float nx = vo.x / ob + (t.dx / tw) * b / ob;
  const float ny = vo.y / ob + (t.dy / tw) * b / ob;
  const float nz = vo.z / ob + (t.dz / tw) * b / ob;
  if (!x_out) {
    v_next[s] = make_float4(nx, ny, nz, 0.0f);
    return;
  }
  const size_t i = (size_t)perm[s];
  x_out[3 * i] = xi.x;
  x_out[3 * i + 1] = xi.y;
  x_out[3 * i + 2] = xi.z;
  v_out[3 * i] = nx;
  v_out[3 * i + 1] = ny;
  v_out[3 * i + 2] = nz;
}

__device__ __forceinline__ VisSums add_parities(const VisSums& e, const VisSums& o) {
  VisSums t;
  t.tw = e.tw + o.tw;
  t.dx = e.dx + o.dx;
  t.dy = e.dy + o.dy;
  t.dz = e.dz + o.dz;
  return t;
}

// One Jacobi iteration (K_ApplyViscosity, ImplicitViscosity.cu:20-62) in
// sorted order over the step's lists, the sums kept as one partial sum per
// entry parity in the neighbor order (as the PBD passes). kPair: thread g
// takes entry parity g & 1 of particle g >> 1 (a lane pair per particle, for
// small searches; waves with 32-bit lists or none run the list on the even
// lane); both forms give the same bits.
template <bool kPair>
__global__ __launch_bounds__(kBlock) void k_viscosity(NbrView nv, const float4* __restrict__ xs,
                                                      const float4* __restrict__ v_old,
                                                      const float4* __restrict__ v_in, float4* __restrict__ v_next,
                                                      VisConst vc, const int* __restrict__ perm,
                                                      float* __restrict__ x_out, float* __restrict__ v_out) {
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int s = kPair ? g >> 1 : g, h = kPair ? g & 1 : 0;
  if (s >= nv.n) return;
  const float4 xi = xs[s];
  auto gather = [&](int j) { return VisPair{xs[j], v_in[j]}; };
  VisSums t;
  if (kPair && for_each_neighbor_half_idx(nv, s, h, gather, [&](int, const VisPair& q) { vis_term(t, xi, q, vc); })) {
    VisSums o;
    o.tw = __shfl_xor(t.tw, 1);
    o.dx = __shfl_xor(t.dx, 1);
    o.dy = __shfl_xor(t.dy, 1);
    o.dz = __shfl_xor(t.dz, 1);
    if (h == 0) vis_finish(s, xi, add_parities(t, o), v_old, v_next, vc, perm, x_out, v_out);
    return;
  }
  if (h != 0) return;
  VisSums a[2];
  for_each_neighbor_p(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)), gather,
      [&](int, const VisPair& q, auto par) { vis_term(a[decltype(par)::value], xi, q, vc); },
      [&](bool odd) {
        const VisSums x = odd ? a[1] : a[0], y = odd ? a[0] : a[1];
        a[0] = x;
        a[1] = y;
      });
  vis_finish(s, xi, add_parities(a[0], a[1]), v_old, v_next, vc, perm, x_out, v_out);
}

void launch_viscosity(const NbrView& nv, const float4* xs, const float4* v_old, const float4* v_in, float4* v_next,
                      const VisConst& vc, const int* perm, float* x_out, float* v_out, hipStream_t st, int lanes) {
  if (lanes == 2)
    k_viscosity<true><<<(int)((2ll * nv.n + kBlock - 1) / kBlock), kBlock, 0, st>>>(nv, xs, v_old, v_in, v_next, vc,
                                                                                    perm, x_out, v_out);
  else
    k_viscosity<false><<<(nv.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nv, xs, v_old, v_in, v_next, vc, perm, x_out,
                                                                        v_out);
}

// Module path: one iteration over a caller CSR list in original order
// (positions xa as float4; v_old, v_in float4; the last iteration writes the
// caller's packed velocities).
__global__ __launch_bounds__(kBlock) void k_csr_viscosity(const float4* __restrict__ xa, Csr nb,
                                                          const float4* __restrict__ v_old,
                                                          const float4* __restrict__ v_in,
                                                          float4* __restrict__ v_next, float* __restrict__ v_out,
                                                          VisConst vc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb.n) return;
  const float4 xi = xa[i];
  float tw = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
  for (int e = nb.begin(i), end = nb.end(i); e < end; ++e) {
    const int j = nb.elems[e];
    const PairGeom g = pair_geom(xi, xa[j], vc.inv_h);
    if (g.r > vc.eps) {
      const float w = vc.cw * g.t;
      const float4 vj = v_in[j];
      tw += w;
      dx += w * vj.x;
      dy += w * vj.y;
      dz += w * vj.z;
    }
  }
  const float b = tw < vc.eps ? 0.0f : vc.b;
  tw = tw < vc.eps ? 1.0f : tw;
  const float ob = 1.0f + b;
  const float4 vo = v_old[i];
  const float nx = vo.x / ob + (dx / tw) * b / ob;
  const float ny = vo.y / ob + (dy / tw) * b / ob;
  const float nz = vo.z / ob + (dz / tw) * b / ob;
  if (v_out) {
    v_out[3 * (size_t)i] = nx;
    v_out[3 * (size_t)i + 1] = ny;
    v_out[3 * (size_t)i + 2] = nz;
  } else {
    v_next[i] = make_float4(nx, ny, nz, 0.0f);
  }
}

void launch_csr_viscosity(const float4* xa, const Csr& nb, const float4* v_old, const float4* v_in, float4* v_next,
                          float* v_out, const VisConst& vc, hipStream_t st) {
  k_csr_viscosity<<<(nb.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, nb, v_old, v_in, v_next, v_out, vc);
}

}  // namespace pbf
