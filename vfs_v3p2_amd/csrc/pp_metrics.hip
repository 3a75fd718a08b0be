// FormMetrics (metrics.c:4-1290) on the device: the face metrics of the i, j and
// k faces with their Jacobians, the cell metrics (NEWMETRIC node-plane averages)
// and the cell Jacobian, from the DMDA node coordinates, in the UserCtx layout
// vfs_pp_set_geometry takes.  One thread per DMDA point; the arithmetic keeps
// the reference's operand order (face centres metrics.c:613-622, derivatives
// 628-700, cross products and Jacobian 700-724; j faces 739-846, k faces
// 859-1004; cell metrics 185-296; cell Jacobian 302-371; mirrored boundary
// planes 374-454 and periodic aliasing 477-509), so the results are bit-exact
// with a CPU restatement compiled without contraction.
#include "pp_engine.hpp"

namespace pp {

namespace {

#pragma clang fp contract(off)

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 ld(const double* __restrict__ co, const Dmda& m, int i, int j, int k) {
  const size_t p = 3 * m.idx(i, j, k);
  return V3{co[p], co[p + 1], co[p + 2]};
}
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// face centre of a face whose four corners are given (sum, then x 0.25)
__device__ __forceinline__ V3 quad(V3 a, V3 b, V3 c, V3 d) {
  return V3{0.25 * (a.x + b.x + c.x + d.x), 0.25 * (a.y + b.y + c.y + d.y), 0.25 * (a.z + b.z + c.z + d.z)};
}

// (csi, eta, zet, 1/J) from the derivative vectors along xi, eta, zeta
__device__ __forceinline__ void cross_metrics(V3 dc, V3 de, V3 dz, V3& csi, V3& eta, V3& zet, double& aj) {
  csi = V3{de.y * dz.z - de.z * dz.y, -de.x * dz.z + de.z * dz.x, de.x * dz.y - de.y * dz.x};
  eta = V3{dz.y * dc.z - dz.z * dc.y, -dz.x * dc.z + dz.z * dc.x, dz.x * dc.y - dz.y * dc.x};
  zet = V3{dc.y * de.z - dc.z * de.y, -dc.x * de.z + dc.z * de.x, dc.x * de.y - dc.y * de.x};
  const double det = dc.x * (de.y * dz.z - de.z * dz.y) - dc.y * (de.x * dz.z - de.z * dz.x) +
                     dc.z * (de.x * dz.y - de.y * dz.x);
  aj = 1.0 / det;
}

// face-centre arrays: c[q](k, j, i) for the i faces (q = 0, corners over j and
// k), the j faces (q = 1, corners over i and k) and the k faces (q = 2, corners
// over i and j); 0 where a corner index would be -1
__device__ __forceinline__ V3 centre(const double* __restrict__ co, const Dmda& m, int q, int i, int j, int k) {
  if (q == 0) {
    if (k < 1 || j < 1) return V3{0., 0., 0.};
    return quad(ld(co, m, i, j, k), ld(co, m, i, j, k - 1), ld(co, m, i, j - 1, k), ld(co, m, i, j - 1, k - 1));
  }
  if (q == 1) {
    if (k < 1 || i < 1) return V3{0., 0., 0.};
    return quad(ld(co, m, i, j, k), ld(co, m, i, j, k - 1), ld(co, m, i - 1, j, k), ld(co, m, i - 1, j, k - 1));
  }
  if (j < 1 || i < 1) return V3{0., 0., 0.};
  return quad(ld(co, m, i, j, k), ld(co, m, i, j - 1, k), ld(co, m, i - 1, j, k), ld(co, m, i - 1, j - 1, k));
}

// derivative of the face-centre array q along axis `ax` at (i, j, k): forward at
// `first`, backward at `last` (two-sided across the seam when periodic), central
// between, 0 outside
__device__ __forceinline__ V3 deriv(const double* __restrict__ co, const Dmda& m, int q, int ax, int first, int last,
                                    bool seam, int i, int j, int k) {
  const int t = ax == 0 ? i : (ax == 1 ? j : k);
  auto at = [&](int s) {
    return ax == 0 ? centre(co, m, q, s, j, k) : (ax == 1 ? centre(co, m, q, i, s, k) : centre(co, m, q, i, j, s));
  };
  if (t == first) return sub(at(first + 1), at(first));
  if (t == last) {
    if (seam) {
      const V3 a = at(1), b = at(0), c = at(last), d = at(last - 1);
      return V3{0.5 * (a.x - b.x + c.x - d.x), 0.5 * (a.y - b.y + c.y - d.y), 0.5 * (a.z - b.z + c.z - d.z)};
    }
    return sub(at(last), at(last - 1));
  }
  if (t > first && t < last) {
    const V3 a = at(t + 1), b = at(t - 1);
    return V3{0.5 * (a.x - b.x), 0.5 * (a.y - b.y), 0.5 * (a.z - b.z)};
  }
  return V3{0., 0., 0.};
}

struct MetricOut {
  double *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
  double *iaj, *jaj, *kaj, *aj, *csi, *eta, *zet;
};

__device__ __forceinline__ void put(double* a, size_t p, V3 v) {
  a[3 * p] = v.x;
  a[3 * p + 1] = v.y;
  a[3 * p + 2] = v.z;
}

// 0.5*(a + b - c - d), the reference's form of a node-plane derivative
__device__ __forceinline__ V3 d4(V3 a, V3 b, V3 c, V3 d) {
  return V3{0.5 * (a.x + b.x - c.x - d.x), 0.5 * (a.y + b.y - c.y - d.y), 0.5 * (a.z + b.z - c.z - d.z)};
}
// 0.5*x + 0.5*y (AxByC, momentum.c:35-40)
__device__ __forceinline__ V3 avg(V3 x, V3 y) {
  return V3{0.5 * x.x + 0.5 * y.x, 0.5 * x.y + 0.5 * y.y, 0.5 * x.z + 0.5 * y.z};
}

// node-plane area vectors of the NEWMETRIC branch (metrics.c:185-270):
// csit(i, j, k) is csitmp of node plane i over nodes j..j+1, k..k+1, etat and
// zett likewise for node planes j and k
__device__ __forceinline__ V3 csit(const double* __restrict__ co, const Dmda& m, int i, int j, int k) {
  const V3 de = d4(ld(co, m, i, j + 1, k + 1), ld(co, m, i, j + 1, k), ld(co, m, i, j, k + 1), ld(co, m, i, j, k));
  const V3 dz = d4(ld(co, m, i, j, k + 1), ld(co, m, i, j + 1, k + 1), ld(co, m, i, j, k), ld(co, m, i, j + 1, k));
  return V3{de.y * dz.z - de.z * dz.y, -de.x * dz.z + de.z * dz.x, de.x * dz.y - de.y * dz.x};
}
__device__ __forceinline__ V3 etat(const double* __restrict__ co, const Dmda& m, int i, int j, int k) {
  const V3 dc = d4(ld(co, m, i + 1, j, k + 1), ld(co, m, i + 1, j, k), ld(co, m, i, j, k + 1), ld(co, m, i, j, k));
  const V3 dz = d4(ld(co, m, i + 1, j, k + 1), ld(co, m, i, j, k + 1), ld(co, m, i + 1, j, k), ld(co, m, i, j, k));
  return V3{dz.y * dc.z - dz.z * dc.y, -dz.x * dc.z + dz.z * dc.x, dz.x * dc.y - dz.y * dc.x};
}
__device__ __forceinline__ V3 zett(const double* __restrict__ co, const Dmda& m, int i, int j, int k) {
  const V3 dc = d4(ld(co, m, i + 1, j + 1, k), ld(co, m, i + 1, j, k), ld(co, m, i, j + 1, k), ld(co, m, i, j, k));
  const V3 de = d4(ld(co, m, i + 1, j + 1, k), ld(co, m, i, j + 1, k), ld(co, m, i + 1, j, k), ld(co, m, i, j, k));
  return V3{dc.y * de.z - dc.z * de.y, -dc.x * de.z + dc.z * de.x, dc.x * de.y - dc.y * de.x};
}

// cell Jacobian from the 8 corners (metrics.c:305-371), interior point
__device__ __forceinline__ double cell_aj(const double* __restrict__ co, const Dmda& m, int i, int j, int k) {
  const V3 c0 = ld(co, m, i, j, k), c1 = ld(co, m, i, j - 1, k), c2 = ld(co, m, i, j, k - 1),
           c3 = ld(co, m, i, j - 1, k - 1), c4 = ld(co, m, i - 1, j, k), c5 = ld(co, m, i - 1, j - 1, k),
           c6 = ld(co, m, i - 1, j, k - 1), c7 = ld(co, m, i - 1, j - 1, k - 1);
  auto comb = [](double a, double b, double c, double d, double e, double f, double g, double h) {
    return 0.25 * (a + b + c + d - e - f - g - h);
  };
  const V3 dC{comb(c0.x, c1.x, c2.x, c3.x, c4.x, c5.x, c6.x, c7.x), comb(c0.y, c1.y, c2.y, c3.y, c4.y, c5.y, c6.y, c7.y),
              comb(c0.z, c1.z, c2.z, c3.z, c4.z, c5.z, c6.z, c7.z)};
  const V3 dE{comb(c0.x, c4.x, c2.x, c6.x, c1.x, c5.x, c3.x, c7.x), comb(c0.y, c4.y, c2.y, c6.y, c1.y, c5.y, c3.y, c7.y),
              comb(c0.z, c4.z, c2.z, c6.z, c1.z, c5.z, c3.z, c7.z)};
  const V3 dZ{comb(c0.x, c1.x, c4.x, c5.x, c2.x, c3.x, c6.x, c7.x), comb(c0.y, c1.y, c4.y, c5.y, c2.y, c3.y, c6.y, c7.y),
              comb(c0.z, c1.z, c4.z, c5.z, c2.z, c3.z, c6.z, c7.z)};
  V3 a, b, c;
  double aj;
  cross_metrics(dC, dE, dZ, a, b, c, aj);
  return aj;
}

// interior index feeding a boundary-plane point: mirrored from the first/last
// interior plane, then replaced by the periodic image on periodic axes
__device__ __forceinline__ int src_index(int t, int mm, int per) {
  if (t == 0) return per ? mm - 2 : 1;
  if (t == mm - 1) return per ? 1 : mm - 2;
  return t;
}

__global__ __launch_bounds__(256) void k_form_metrics(Dmda m, const double* __restrict__ co, MetricOut o) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= m.mx || j >= m.my) return;
  const size_t p = m.idx(i, j, k);
  const int mx = m.mx, my = m.my, mz = m.mz;
  // i faces (metrics.c:613-724): points k in [1, mz-2], j in [1, my-2], i in [0, mx-2]
  {
    V3 a{0., 0., 0.}, b = a, c = a;
    double aj = 0.;
    if (k >= 1 && k <= mz - 2 && j >= 1 && j <= my - 2 && i <= mx - 2)
      cross_metrics(deriv(co, m, 0, 0, 0, mx - 2, m.per[0], i, j, k), deriv(co, m, 0, 1, 1, my - 2, false, i, j, k),
                    deriv(co, m, 0, 2, 1, mz - 2, false, i, j, k), a, b, c, aj);
    put(o.icsi, p, a);
    put(o.ieta, p, b);
    put(o.izet, p, c);
    o.iaj[p] = aj;
  }
  // j faces (metrics.c:739-846): k in [1, mz-2], j in [0, my-2], i in [1, mx-2]
  {
    V3 a{0., 0., 0.}, b = a, c = a;
    double aj = 0.;
    if (k >= 1 && k <= mz - 2 && j <= my - 2 && i >= 1 && i <= mx - 2)
      cross_metrics(deriv(co, m, 1, 0, 1, mx - 2, false, i, j, k), deriv(co, m, 1, 1, 0, my - 2, m.per[1], i, j, k),
                    deriv(co, m, 1, 2, 1, mz - 2, false, i, j, k), a, b, c, aj);
    put(o.jcsi, p, a);
    put(o.jeta, p, b);
    put(o.jzet, p, c);
    o.jaj[p] = aj;
  }
  // k faces (metrics.c:859-1004): k in [0, mz-2], j in [1, my-2], i in [1, mx-2]
  {
    V3 a{0., 0., 0.}, b = a, c = a;
    double aj = 0.;
    if (k <= mz - 2 && j >= 1 && j <= my - 2 && i >= 1 && i <= mx - 2)
      cross_metrics(deriv(co, m, 2, 0, 1, mx - 2, false, i, j, k), deriv(co, m, 2, 1, 1, my - 2, false, i, j, k),
                    deriv(co, m, 2, 2, 0, mz - 2, m.per[2], i, j, k), a, b, c, aj);
    put(o.kcsi, p, a);
    put(o.keta, p, b);
    put(o.kzet, p, c);
    o.kaj[p] = aj;
  }
  // cell metrics and Jacobian at the interior point this one takes its values from
  const int si = src_index(i, mx, m.per[0]), sj = src_index(j, my, m.per[1]), sk = src_index(k, mz, m.per[2]);
  put(o.csi, p, avg(csit(co, m, si, sj - 1, sk - 1), csit(co, m, si - 1, sj - 1, sk - 1)));
  put(o.eta, p, avg(etat(co, m, si - 1, sj, sk - 1), etat(co, m, si - 1, sj - 1, sk - 1)));
  put(o.zet, p, avg(zett(co, m, si - 1, sj - 1, sk), zett(co, m, si - 1, sj - 1, sk - 1)));
  o.aj[p] = cell_aj(co, m, si, sj, sk);
}

}  // namespace

void launch_form_metrics(hipStream_t s, Dmda m, const double* coor, double* const out[16]) {
  MetricOut o{out[0], out[1], out[2], out[3], out[4], out[5], out[6], out[7], out[8],
              out[9], out[10], out[11], out[12], out[13], out[14], out[15]};
  k_form_metrics<<<dim3((m.mx + 63) / 64, (m.my + 3) / 4, m.mz), dim3(64, 4, 1), 0, s>>>(m, coor, o);
  HIPCHECK(hipGetLastError());
}

}  // namespace pp
