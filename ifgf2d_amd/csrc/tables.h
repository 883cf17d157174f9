// Per-cone geometry records and Chebyshev tables shared by the host plan code
// and the gfx950 kernels.
#pragma once

namespace ifgf {

struct ConeGeo {
  double cx, cy;  // box centre (Level::GetBoxCenter)
  double h;       // Level::scaling_r2s_
  double rmin;    // h / radial[posr + 1]
  double dr;      // h / radial[posr] - rmin
  double da;      // pi / nconeselevation_
  double th0;     // da * postheta
  double pad;
};

// Chebyshev tables of the cone interpolant (Chebyshev.h:10-68), up to 16x16,
// in FP64 (plain plans) or rounded to FP32 (the cone transforms of a
// mixed-precision plan).
template <typename R>
struct ChebTablesT {
  R xs[16], xa[16];    // ascending Chebyshev points in s and theta
  R Ts[16 * 16];       // Ts[m*16 + k] = T_m(xs[k]) * alpha_m
  R Ta[16 * 16];       // Ta[n*16 + l] = T_n(xa[l]) * alpha_n / (Ps*Pang)
};
using ChebTables = ChebTablesT<double>;

// Copies the transform rows a (Ps, Pang) kernel uses -- Ts[m][k], m, k < Ps,
// and Ta[n][l], n, l < Pang -- from src to the block's LDS copy dst (all
// threads of the block call it; a barrier must follow before dst is read).
template <int Ps, int Pang, typename R>
__device__ inline void load_transform_rows(ChebTablesT<R>& dst, const ChebTablesT<R>* __restrict__ src) {
  for (int q = threadIdx.x; q < Ps * Ps + Pang * Pang; q += blockDim.x) {
    if (q < Ps * Ps) {
      const int o = (q / Ps) * 16 + q % Ps;
      dst.Ts[o] = src->Ts[o];
    } else {
      const int r = q - Ps * Ps;
      const int o = (r / Pang) * 16 + r % Pang;
      dst.Ta[o] = src->Ta[o];
    }
  }
}
using ChebTablesF = ChebTablesT<float>;

// One near-field work item: kNearItemChunks (or fewer) chunks of one target
// group, starting at value offset off.
struct NearItem {
  long long off;
  int g;
  short u0, nc;
};

// One leaf-operator work item: up to kLeafItemTiles tiles of one leaf cone;
// off = value offset of its first tile, s0 = density index of its first
// source, n = sources in the item.
struct LeafItem {
  long long off;
  int s0, n;
};

// One level-push task: cones [c0, c0 + nc) of one level (global ids) and
// their push entries [e0, e0 + ne).
struct PushTask {
  int c0, nc;
  int e0, ne;
};

// 8-column MFMA blocks (kernels_batched.h). A group is a leaf box: its
// targets (near field) or its leaf cones' (cone, node) rows (leaf operator)
// against its sources; aoff = offset of its A tiles, rows, nsrc = sources,
// row0 = first output row, seg = the near group's segment table index (the
// leaf group's first source point). A unit is one row tile rt of group g.
struct MfmaGroup {
  long long aoff;
  int rows, nsrc, row0, seg;
};
struct MfmaUnit {
  int g, rt;
};

// Local Chebyshev coordinates (s, theta) of a point in a cone and the centre
// factor exp(i k r)/sqrt(r) at its distance r from the cone's box centre
// (evaluated in FP64; stored rounded to FP32 on a mixed-precision plan).
template <typename R>
struct Geo4T {
  R s, t;
  R fre, fim;
};
using Geo4 = Geo4T<double>;
using Geo4F = Geo4T<float>;

}  // namespace ifgf
