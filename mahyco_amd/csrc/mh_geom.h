// mh_geom.h -- corner vectors (CQS), volumes and small vector helpers shared by the
// gfx950 kernels and the host-side case initialisation.
//
// Every formula keeps the reference's operation order so the device reproduces the
// CPU reference bit for bit (no FMA contraction: compiled with -ffp-contract=off):
//   hexahedron CQS       MahycoModule::computeCQs, MahycoModule.cc:1869-1932
//   face centres         MahycoModule::computeGeometricValues, MahycoModule.cc:1120-1125
//   quadrangle CQS       MahycoModule.cc:1127-1146
//   volume               MahycoModule.cc:1149-1162  (V = sum_n x_n . CQS_n / dim)
#ifndef MH_GEOM_H
#define MH_GEOM_H

#include "mh_math.h"

struct D3 {
  double x, y, z;
};
MH_HD D3 d3(double x, double y, double z) { D3 r; r.x = x; r.y = y; r.z = z; return r; }
MH_HD D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
MH_HD D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
MH_HD D3 operator*(double s, D3 a) { return d3(s * a.x, s * a.y, s * a.z); }
MH_HD double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// half cross product 0.5 * vecMul(a, b)
MH_HD D3 hcross(D3 a, D3 b) {
  return 0.5 * d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Face centres of a hexahedron, local face order of Arcane's Hexaedron8.
MH_HD D3 hex_fc0(const D3* x) { return 0.25 * (x[0] + x[3] + x[2] + x[1]); }
MH_HD D3 hex_fc1(const D3* x) { return 0.25 * (x[0] + x[4] + x[7] + x[3]); }
MH_HD D3 hex_fc2(const D3* x) { return 0.25 * (x[0] + x[1] + x[5] + x[4]); }
MH_HD D3 hex_fc3(const D3* x) { return 0.25 * (x[4] + x[5] + x[6] + x[7]); }
MH_HD D3 hex_fc4(const D3* x) { return 0.25 * (x[1] + x[2] + x[6] + x[5]); }
MH_HD D3 hex_fc5(const D3* x) { return 0.25 * (x[2] + x[3] + x[7] + x[6]); }

// All eight corner vectors of a hexahedron.
MH_HD void cqs_hex(const D3* x, D3* q) {
  const D3 c0 = hex_fc0(x), c1 = hex_fc1(x), c2 = hex_fc2(x);
  const D3 c3 = hex_fc3(x), c4 = hex_fc4(x), c5 = hex_fc5(x);
  const D3 n1a04 = hcross(x[0] - c0, x[3] - c0), n1a03 = hcross(x[3] - c0, x[2] - c0);
  const D3 n1a02 = hcross(x[2] - c0, x[1] - c0), n1a01 = hcross(x[1] - c0, x[0] - c0);
  const D3 n2a05 = hcross(x[0] - c1, x[4] - c1), n2a12 = hcross(x[4] - c1, x[7] - c1);
  const D3 n2a08 = hcross(x[7] - c1, x[3] - c1), n2a04 = hcross(x[3] - c1, x[0] - c1);
  const D3 n3a01 = hcross(x[0] - c2, x[1] - c2), n3a06 = hcross(x[1] - c2, x[5] - c2);
  const D3 n3a09 = hcross(x[5] - c2, x[4] - c2), n3a05 = hcross(x[4] - c2, x[0] - c2);
  const D3 n4a09 = hcross(x[4] - c3, x[5] - c3), n4a10 = hcross(x[5] - c3, x[6] - c3);
  const D3 n4a11 = hcross(x[6] - c3, x[7] - c3), n4a12 = hcross(x[7] - c3, x[4] - c3);
  const D3 n5a02 = hcross(x[1] - c4, x[2] - c4), n5a07 = hcross(x[2] - c4, x[6] - c4);
  const D3 n5a10 = hcross(x[6] - c4, x[5] - c4), n5a06 = hcross(x[5] - c4, x[1] - c4);
  const D3 n6a03 = hcross(x[2] - c5, x[3] - c5), n6a08 = hcross(x[3] - c5, x[7] - c5);
  const D3 n6a11 = hcross(x[7] - c5, x[6] - c5), n6a07 = hcross(x[6] - c5, x[2] - c5);
  const double t = 1. / 12.;
  q[0] = t * (5. * (n1a01 + n1a04 + n2a04 + n2a05 + n3a05 + n3a01) + (n1a02 + n1a03 + n2a08 + n2a12 + n3a06 + n3a09));
  q[1] = t * (5. * (n1a01 + n1a02 + n3a01 + n3a06 + n5a06 + n5a02) + (n1a04 + n1a03 + n3a09 + n3a05 + n5a10 + n5a07));
  q[2] = t * (5. * (n1a02 + n1a03 + n5a07 + n5a02 + n6a07 + n6a03) + (n1a01 + n1a04 + n5a06 + n5a10 + n6a11 + n6a08));
  q[3] = t * (5. * (n1a03 + n1a04 + n2a08 + n2a04 + n6a08 + n6a03) + (n1a01 + n1a02 + n2a05 + n2a12 + n6a07 + n6a11));
  q[4] = t * (5. * (n2a05 + n2a12 + n3a05 + n3a09 + n4a09 + n4a12) + (n2a08 + n2a04 + n3a01 + n3a06 + n4a10 + n4a11));
  q[5] = t * (5. * (n3a06 + n3a09 + n4a09 + n4a10 + n5a10 + n5a06) + (n3a01 + n3a05 + n4a12 + n4a11 + n5a07 + n5a02));
  q[6] = t * (5. * (n4a11 + n4a10 + n5a10 + n5a07 + n6a07 + n6a11) + (n4a12 + n4a09 + n5a06 + n5a02 + n6a03 + n6a08));
  q[7] = t * (5. * (n2a08 + n2a12 + n4a12 + n4a11 + n6a11 + n6a08) + (n2a04 + n2a05 + n4a09 + n4a10 + n6a07 + n6a03));
}

// All eight corner vectors of a hexahedron, handed to f(l, q_l) as soon as each is
// complete.  Faces are formed in the order 1, 2, 3, 5, 6, 4 so that face 1 (corners 0..3)
// retires before face 4 (corners 4..7) is formed: five faces are live at most instead of
// six.  Every term and sum is the one of cqs_hex, so each q_l has the same bits.
template <class F>
MH_HD void cqs_hex_each(const D3* x, F&& f) {
  const double t = 1. / 12.;
  const D3 c0 = hex_fc0(x);
  const D3 n1a04 = hcross(x[0] - c0, x[3] - c0), n1a03 = hcross(x[3] - c0, x[2] - c0);
  const D3 n1a02 = hcross(x[2] - c0, x[1] - c0), n1a01 = hcross(x[1] - c0, x[0] - c0);
  const D3 c1 = hex_fc1(x);
  const D3 n2a05 = hcross(x[0] - c1, x[4] - c1), n2a12 = hcross(x[4] - c1, x[7] - c1);
  const D3 n2a08 = hcross(x[7] - c1, x[3] - c1), n2a04 = hcross(x[3] - c1, x[0] - c1);
  const D3 c2 = hex_fc2(x);
  const D3 n3a01 = hcross(x[0] - c2, x[1] - c2), n3a06 = hcross(x[1] - c2, x[5] - c2);
  const D3 n3a09 = hcross(x[5] - c2, x[4] - c2), n3a05 = hcross(x[4] - c2, x[0] - c2);
  f(0, t * (5. * (n1a01 + n1a04 + n2a04 + n2a05 + n3a05 + n3a01) + (n1a02 + n1a03 + n2a08 + n2a12 + n3a06 + n3a09)));
  const D3 c4 = hex_fc4(x);
  const D3 n5a02 = hcross(x[1] - c4, x[2] - c4), n5a07 = hcross(x[2] - c4, x[6] - c4);
  const D3 n5a10 = hcross(x[6] - c4, x[5] - c4), n5a06 = hcross(x[5] - c4, x[1] - c4);
  f(1, t * (5. * (n1a01 + n1a02 + n3a01 + n3a06 + n5a06 + n5a02) + (n1a04 + n1a03 + n3a09 + n3a05 + n5a10 + n5a07)));
  const D3 c5 = hex_fc5(x);
  const D3 n6a03 = hcross(x[2] - c5, x[3] - c5), n6a08 = hcross(x[3] - c5, x[7] - c5);
  const D3 n6a11 = hcross(x[7] - c5, x[6] - c5), n6a07 = hcross(x[6] - c5, x[2] - c5);
  f(2, t * (5. * (n1a02 + n1a03 + n5a07 + n5a02 + n6a07 + n6a03) + (n1a01 + n1a04 + n5a06 + n5a10 + n6a11 + n6a08)));
  f(3, t * (5. * (n1a03 + n1a04 + n2a08 + n2a04 + n6a08 + n6a03) + (n1a01 + n1a02 + n2a05 + n2a12 + n6a07 + n6a11)));
  const D3 c3 = hex_fc3(x);
  const D3 n4a09 = hcross(x[4] - c3, x[5] - c3), n4a10 = hcross(x[5] - c3, x[6] - c3);
  const D3 n4a11 = hcross(x[6] - c3, x[7] - c3), n4a12 = hcross(x[7] - c3, x[4] - c3);
  f(4, t * (5. * (n2a05 + n2a12 + n3a05 + n3a09 + n4a09 + n4a12) + (n2a08 + n2a04 + n3a01 + n3a06 + n4a10 + n4a11)));
  f(5, t * (5. * (n3a06 + n3a09 + n4a09 + n4a10 + n5a10 + n5a06) + (n3a01 + n3a05 + n4a12 + n4a11 + n5a07 + n5a02)));
  f(6, t * (5. * (n4a11 + n4a10 + n5a10 + n5a07 + n6a07 + n6a11) + (n4a12 + n4a09 + n5a06 + n5a02 + n6a03 + n6a08)));
  f(7, t * (5. * (n2a08 + n2a12 + n4a12 + n4a11 + n6a11 + n6a08) + (n2a04 + n2a05 + n4a09 + n4a10 + n6a07 + n6a03)));
}

// One corner vector of a hexahedron: only the three faces around corner L are formed;
// each term is computed exactly as in cqs_hex, so the result has the same bits.
template <int L>
MH_HD D3 cqs_hex_corner(const D3* x) {
  constexpr bool F1 = (L <= 3);
  constexpr bool F2 = (L == 0 || L == 3 || L == 4 || L == 7);
  constexpr bool F3 = (L == 0 || L == 1 || L == 4 || L == 5);
  constexpr bool F4 = (L >= 4);
  constexpr bool F5 = (L == 1 || L == 2 || L == 5 || L == 6);
  constexpr bool F6 = (L == 2 || L == 3 || L == 6 || L == 7);
  D3 n1a04, n1a03, n1a02, n1a01, n2a05, n2a12, n2a08, n2a04, n3a01, n3a06, n3a09, n3a05;
  D3 n4a09, n4a10, n4a11, n4a12, n5a02, n5a07, n5a10, n5a06, n6a03, n6a08, n6a11, n6a07;
  if (F1) {
    const D3 c0 = hex_fc0(x);
    n1a04 = hcross(x[0] - c0, x[3] - c0); n1a03 = hcross(x[3] - c0, x[2] - c0);
    n1a02 = hcross(x[2] - c0, x[1] - c0); n1a01 = hcross(x[1] - c0, x[0] - c0);
  }
  if (F2) {
    const D3 c1 = hex_fc1(x);
    n2a05 = hcross(x[0] - c1, x[4] - c1); n2a12 = hcross(x[4] - c1, x[7] - c1);
    n2a08 = hcross(x[7] - c1, x[3] - c1); n2a04 = hcross(x[3] - c1, x[0] - c1);
  }
  if (F3) {
    const D3 c2 = hex_fc2(x);
    n3a01 = hcross(x[0] - c2, x[1] - c2); n3a06 = hcross(x[1] - c2, x[5] - c2);
    n3a09 = hcross(x[5] - c2, x[4] - c2); n3a05 = hcross(x[4] - c2, x[0] - c2);
  }
  if (F4) {
    const D3 c3 = hex_fc3(x);
    n4a09 = hcross(x[4] - c3, x[5] - c3); n4a10 = hcross(x[5] - c3, x[6] - c3);
    n4a11 = hcross(x[6] - c3, x[7] - c3); n4a12 = hcross(x[7] - c3, x[4] - c3);
  }
  if (F5) {
    const D3 c4 = hex_fc4(x);
    n5a02 = hcross(x[1] - c4, x[2] - c4); n5a07 = hcross(x[2] - c4, x[6] - c4);
    n5a10 = hcross(x[6] - c4, x[5] - c4); n5a06 = hcross(x[5] - c4, x[1] - c4);
  }
  if (F6) {
    const D3 c5 = hex_fc5(x);
    n6a03 = hcross(x[2] - c5, x[3] - c5); n6a08 = hcross(x[3] - c5, x[7] - c5);
    n6a11 = hcross(x[7] - c5, x[6] - c5); n6a07 = hcross(x[6] - c5, x[2] - c5);
  }
  const double t = 1. / 12.;
  if (L == 0) return t * (5. * (n1a01 + n1a04 + n2a04 + n2a05 + n3a05 + n3a01) + (n1a02 + n1a03 + n2a08 + n2a12 + n3a06 + n3a09));
  if (L == 1) return t * (5. * (n1a01 + n1a02 + n3a01 + n3a06 + n5a06 + n5a02) + (n1a04 + n1a03 + n3a09 + n3a05 + n5a10 + n5a07));
  if (L == 2) return t * (5. * (n1a02 + n1a03 + n5a07 + n5a02 + n6a07 + n6a03) + (n1a01 + n1a04 + n5a06 + n5a10 + n6a11 + n6a08));
  if (L == 3) return t * (5. * (n1a03 + n1a04 + n2a08 + n2a04 + n6a08 + n6a03) + (n1a01 + n1a02 + n2a05 + n2a12 + n6a07 + n6a11));
  if (L == 4) return t * (5. * (n2a05 + n2a12 + n3a05 + n3a09 + n4a09 + n4a12) + (n2a08 + n2a04 + n3a01 + n3a06 + n4a10 + n4a11));
  if (L == 5) return t * (5. * (n3a06 + n3a09 + n4a09 + n4a10 + n5a10 + n5a06) + (n3a01 + n3a05 + n4a12 + n4a11 + n5a07 + n5a02));
  if (L == 6) return t * (5. * (n4a11 + n4a10 + n5a10 + n5a07 + n6a07 + n6a11) + (n4a12 + n4a09 + n5a06 + n5a02 + n6a03 + n6a08));
  return t * (5. * (n2a08 + n2a12 + n4a12 + n4a11 + n6a11 + n6a08) + (n2a04 + n2a05 + n4a09 + n4a10 + n6a07 + n6a03));
}

// Component C (0 = x, 1 = y, 2 = z) of the eight corner vectors of a hexahedron, with the
// operations of cqs_hex on that component only: the face centre components, the differences
// to the centre, the component of each half cross product 0.5 (a_U b_W - a_W b_U) (U = C+1,
// W = C+2 mod 3) and the corner sums t (5 (..) + (..)) in the same order.  Every rounded
// operation of cqs_hex acts on one component, so q[l] is bit for bit cqs_hex's q[l].C, and
// only one component's terms are live at a time.
template <int C>
MH_HD void cqs_hex_comp(const D3* x, double* q) {
  constexpr int U = (C + 1) % 3, W = (C + 2) % 3;
  auto cmp = [](const D3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); };
  // one face with nodes p0..p3 in the order of its centre sum: centre, then the cross
  // products of the edges p0p1, p1p2, p2p3, p3p0 (names of cqs_hex)
  auto face = [&](int p0, int p1, int p2, int p3, double& e0, double& e1, double& e2, double& e3) {
    const double cu = 0.25 * (((cmp(x[p0], U) + cmp(x[p1], U)) + cmp(x[p2], U)) + cmp(x[p3], U));
    const double cw = 0.25 * (((cmp(x[p0], W) + cmp(x[p1], W)) + cmp(x[p2], W)) + cmp(x[p3], W));
    const double au0 = cmp(x[p0], U) - cu, aw0 = cmp(x[p0], W) - cw;
    const double au1 = cmp(x[p1], U) - cu, aw1 = cmp(x[p1], W) - cw;
    const double au2 = cmp(x[p2], U) - cu, aw2 = cmp(x[p2], W) - cw;
    const double au3 = cmp(x[p3], U) - cu, aw3 = cmp(x[p3], W) - cw;
    e0 = 0.5 * (au0 * aw1 - aw0 * au1);
    e1 = 0.5 * (au1 * aw2 - aw1 * au2);
    e2 = 0.5 * (au2 * aw3 - aw2 * au3);
    e3 = 0.5 * (au3 * aw0 - aw3 * au0);
  };
  double n1a04, n1a03, n1a02, n1a01, n2a05, n2a12, n2a08, n2a04, n3a01, n3a06, n3a09, n3a05;
  double n4a09, n4a10, n4a11, n4a12, n5a02, n5a07, n5a10, n5a06, n6a03, n6a08, n6a11, n6a07;
  face(0, 3, 2, 1, n1a04, n1a03, n1a02, n1a01);
  face(0, 4, 7, 3, n2a05, n2a12, n2a08, n2a04);
  face(0, 1, 5, 4, n3a01, n3a06, n3a09, n3a05);
  face(4, 5, 6, 7, n4a09, n4a10, n4a11, n4a12);
  face(1, 2, 6, 5, n5a02, n5a07, n5a10, n5a06);
  face(2, 3, 7, 6, n6a03, n6a08, n6a11, n6a07);
  const double t = 1. / 12.;
  q[0] = t * (5. * (n1a01 + n1a04 + n2a04 + n2a05 + n3a05 + n3a01) + (n1a02 + n1a03 + n2a08 + n2a12 + n3a06 + n3a09));
  q[1] = t * (5. * (n1a01 + n1a02 + n3a01 + n3a06 + n5a06 + n5a02) + (n1a04 + n1a03 + n3a09 + n3a05 + n5a10 + n5a07));
  q[2] = t * (5. * (n1a02 + n1a03 + n5a07 + n5a02 + n6a07 + n6a03) + (n1a01 + n1a04 + n5a06 + n5a10 + n6a11 + n6a08));
  q[3] = t * (5. * (n1a03 + n1a04 + n2a08 + n2a04 + n6a08 + n6a03) + (n1a01 + n1a02 + n2a05 + n2a12 + n6a07 + n6a11));
  q[4] = t * (5. * (n2a05 + n2a12 + n3a05 + n3a09 + n4a09 + n4a12) + (n2a08 + n2a04 + n3a01 + n3a06 + n4a10 + n4a11));
  q[5] = t * (5. * (n3a06 + n3a09 + n4a09 + n4a10 + n5a10 + n5a06) + (n3a01 + n3a05 + n4a12 + n4a11 + n5a07 + n5a02));
  q[6] = t * (5. * (n4a11 + n4a10 + n5a10 + n5a07 + n6a07 + n6a11) + (n4a12 + n4a09 + n5a06 + n5a02 + n6a03 + n6a08));
  q[7] = t * (5. * (n2a08 + n2a12 + n4a12 + n4a11 + n6a11 + n6a08) + (n2a04 + n2a05 + n4a09 + n4a10 + n6a07 + n6a03));
}

// cqs_hex_comp<C> with the corner sums accumulated face by face.  Every corner sum of cqs_hex
// adds two terms of each of its three faces, and its faces come in increasing order (corner 0:
// faces 1, 2, 3; 1: 1, 3, 5; 2: 1, 5, 6; 3: 1, 2, 6; 4: 2, 3, 4; 5: 3, 4, 5; 6: 4, 5, 6; 7: 2,
// 4, 6), so running the faces 1..6 and adding each face's terms to the sums of its four corners
// performs every sum's additions in cqs_hex's order: same bits.  Only the 16 running sums and one
// face's four terms are live, instead of the 24 terms of all faces.
template <int C>
MH_HD void cqs_hex_comp_fm(const D3* x, double* q) {
  constexpr int U = (C + 1) % 3, W = (C + 2) % 3;
  auto cmp = [](const D3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); };
  auto face = [&](int p0, int p1, int p2, int p3, double& e0, double& e1, double& e2, double& e3) {
    const double cu = 0.25 * (((cmp(x[p0], U) + cmp(x[p1], U)) + cmp(x[p2], U)) + cmp(x[p3], U));
    const double cw = 0.25 * (((cmp(x[p0], W) + cmp(x[p1], W)) + cmp(x[p2], W)) + cmp(x[p3], W));
    const double au0 = cmp(x[p0], U) - cu, aw0 = cmp(x[p0], W) - cw;
    const double au1 = cmp(x[p1], U) - cu, aw1 = cmp(x[p1], W) - cw;
    const double au2 = cmp(x[p2], U) - cu, aw2 = cmp(x[p2], W) - cw;
    const double au3 = cmp(x[p3], U) - cu, aw3 = cmp(x[p3], W) - cw;
    e0 = 0.5 * (au0 * aw1 - aw0 * au1);
    e1 = 0.5 * (au1 * aw2 - aw1 * au2);
    e2 = 0.5 * (au2 * aw3 - aw2 * au3);
    e3 = 0.5 * (au3 * aw0 - aw3 * au0);
  };
  const double t = 1. / 12.;
  double A[8], B[8];  // the 5 x part and the 1 x part of each corner sum
  double a, b, c, d;
  face(0, 3, 2, 1, a, b, c, d);  // n1a04, n1a03, n1a02, n1a01
  A[0] = d + a; B[0] = c + b;
  A[1] = d + c; B[1] = a + b;
  A[2] = c + b; B[2] = d + a;
  A[3] = b + a; B[3] = d + c;
  face(0, 4, 7, 3, a, b, c, d);  // n2a05, n2a12, n2a08, n2a04
  A[0] = (A[0] + d) + a; B[0] = (B[0] + c) + b;
  A[3] = (A[3] + c) + d; B[3] = (B[3] + a) + b;
  A[4] = a + b; B[4] = c + d;
  A[7] = c + b; B[7] = d + a;
  face(0, 1, 5, 4, a, b, c, d);  // n3a01, n3a06, n3a09, n3a05
  A[0] = (A[0] + d) + a; B[0] = (B[0] + b) + c;
  A[1] = (A[1] + a) + b; B[1] = (B[1] + c) + d;
  A[4] = (A[4] + d) + c; B[4] = (B[4] + a) + b;
  A[5] = b + c; B[5] = a + d;
  q[0] = t * (5. * A[0] + B[0]);
  face(4, 5, 6, 7, a, b, c, d);  // n4a09, n4a10, n4a11, n4a12
  A[4] = (A[4] + a) + d; B[4] = (B[4] + b) + c;
  A[5] = (A[5] + a) + b; B[5] = (B[5] + d) + c;
  A[6] = c + b; B[6] = d + a;
  A[7] = (A[7] + d) + c; B[7] = (B[7] + a) + b;
  q[4] = t * (5. * A[4] + B[4]);
  face(1, 2, 6, 5, a, b, c, d);  // n5a02, n5a07, n5a10, n5a06
  A[1] = (A[1] + d) + a; B[1] = (B[1] + c) + b;
  A[2] = (A[2] + b) + a; B[2] = (B[2] + d) + c;
  A[5] = (A[5] + c) + d; B[5] = (B[5] + b) + a;
  A[6] = (A[6] + c) + b; B[6] = (B[6] + d) + a;
  q[1] = t * (5. * A[1] + B[1]);
  q[5] = t * (5. * A[5] + B[5]);
  face(2, 3, 7, 6, a, b, c, d);  // n6a03, n6a08, n6a11, n6a07
  A[2] = (A[2] + d) + a; B[2] = (B[2] + c) + b;
  A[3] = (A[3] + b) + a; B[3] = (B[3] + d) + c;
  A[6] = (A[6] + d) + c; B[6] = (B[6] + a) + b;
  A[7] = (A[7] + c) + b; B[7] = (B[7] + d) + a;
  q[2] = t * (5. * A[2] + B[2]);
  q[3] = t * (5. * A[3] + B[3]);
  q[6] = t * (5. * A[6] + B[6]);
  q[7] = t * (5. * A[7] + B[7]);
}

// Volume of a hexahedron, sum_l x_l . CQS_l / 3 (MahycoModule.cc:1149-1162), one corner-vector
// component at a time: dot3 rounds (x.x q.x + x.y q.y) + x.z q.z, so the products of the x and
// y passes are summed before the z pass, and the corner terms are added in corner order.  Same
// bits as cell_volume<8>(x, q) after cqs_hex(x, q).
MH_HD double hex_volume_lowreg(const D3* x) {
  double q[8], d[8];
  cqs_hex_comp<0>(x, q);
  for (int l = 0; l < 8; ++l) d[l] = x[l].x * q[l];
  cqs_hex_comp<1>(x, q);
  for (int l = 0; l < 8; ++l) d[l] = d[l] + x[l].y * q[l];
  cqs_hex_comp<2>(x, q);
  double v = 0.;
  for (int l = 0; l < 8; ++l) v += d[l] + x[l].z * q[l];
  return v / 3;
}

// Quadrangle corner vectors: half edge normals of the two edges at each corner
// (x[4] = x[0], npc[0] = npc[4]; cqs[ii] = npc[ii+1] + npc[ii]).
MH_HD D3 quad_npc(const D3* x, int k) {  // npc[k+1], k = 0..3
  const D3 a = x[k], b = x[(k + 1) & 3];
  return d3(0.5 * (b.y - a.y), 0.5 * (a.x - b.x), 0.);
}
template <int L>
MH_HD D3 cqs_quad_corner(const D3* x) {
  const D3 p1 = quad_npc(x, L);            // npc[L+1]
  const D3 p0 = quad_npc(x, (L + 3) & 3);  // npc[L] (npc[0] = npc[4])
  return p1 + p0;
}
MH_HD void cqs_quad(const D3* x, D3* q) {
  q[0] = cqs_quad_corner<0>(x);
  q[1] = cqs_quad_corner<1>(x);
  q[2] = cqs_quad_corner<2>(x);
  q[3] = cqs_quad_corner<3>(x);
}

// V = sum_n x_n . CQS_n / dim
template <int NNODE>
MH_HD double cell_volume(const D3* x, const D3* q, int dim) {
  double v = 0.;
  for (int l = 0; l < NNODE; ++l) v += dot3(x[l], q[l]);
  return v / dim;
}

// RemapADIService::fluxLimiter, UtilesRemap.cc:108-120 (classical limiters).
MH_HD double mh_flux_limiter(int id, double r) {
  // std::min(a, b) = (b < a) ? b : a ; std::max(a, b) = (a < b) ? b : a
  if (id == 0) {
    double t = (r < 1.0) ? r : 1.0;  // std::min(1.0, r)
    return (0.0 < t) ? t : 0.0;      // std::max(0.0, t)
  }
  if (id == 1) {
    double a = (1.0 < 2.0 * r) ? 1.0 : 2.0 * r;  // std::min(2r, 1)
    double b = (2.0 < r) ? 2.0 : r;              // std::min(r, 2)
    double c = (a < b) ? b : a;                  // std::max(a, b)
    return (0.0 < c) ? c : 0.0;                  // std::max(0, c)
  }
  if (id == 2) {
    if (r <= 0.0) return 0.0;
    return 2.0 * r / (1.0 + r);
  }
  return 0.0;
}

// LIM value of kernels that read the limiter id at run time and also take the generalised
// limiters (ids >= MH_MINMOD_G, handled by the caller).
#define MH_LIM_ANY 8

// The limiter with its id known at compile time (LIM >= 0: the branches on the id fold
// away) or read at run time (LIM < 0, batched launches over cases with different limiters).
template <int LIM>
MH_HD double mh_limiter(int id, double r) {
  return mh_flux_limiter((LIM < 0 || LIM == MH_LIM_ANY) ? id : LIM, r);
}

#endif
