// Host-side index space: IntVect / Box / dense IntVectSet / DisjointBoxLayout.
//
// Semantics follow Chombo 3.2's BoxTools: floor coarsening
// (lib/src/BoxTools/IntVect.H:1288-1299), Box::operator< ordering of a closed
// DisjointBoxLayout (lib/src/BoxTools/Box.H:2223, BoxLayout.cpp:89-117), and the
// DenseIntVectSet operations used by the coarse-fine machinery
// (lib/src/BoxTools/DenseIntVectSet.cpp).  Only index math lives here; field data
// is on the device.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace amrmg {

using IV = std::array<int, 3>;

inline int floordiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

inline bool lexLT(const IV& a, const IV& b) {
  for (int d = 0; d < 3; d++) {
    if (a[d] < b[d]) return true;
    if (a[d] > b[d]) return false;
  }
  return false;
}

struct Box {
  IV lo{0, 0, 0}, hi{-1, -1, -1};
  Box() = default;
  Box(IV l, IV h) : lo(l), hi(h) {}
  bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
  int len(int d) const { return hi[d] - lo[d] + 1; }
  long long npts() const { return empty() ? 0 : (long long)len(0) * len(1) * len(2); }
  bool contains(int i, int j, int k) const {
    return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
  }
  bool contains(const IV& p) const { return contains(p[0], p[1], p[2]); }
  bool contains(const Box& b) const {
    for (int d = 0; d < 3; d++)
      if (b.lo[d] < lo[d] || b.hi[d] > hi[d]) return false;
    return true;
  }
  bool operator==(const Box& o) const { return lo == o.lo && hi == o.hi; }
  bool operator<(const Box& o) const {
    if (lo == o.lo) return lexLT(hi, o.hi);
    return lexLT(lo, o.lo);
  }
};

inline Box grow(Box b, int g) {
  for (int d = 0; d < 3; d++) { b.lo[d] -= g; b.hi[d] += g; }
  return b;
}
inline Box growDir(Box b, int d, int g) { b.lo[d] -= g; b.hi[d] += g; return b; }
inline Box operator&(Box a, const Box& b) {
  for (int d = 0; d < 3; d++) { a.lo[d] = std::max(a.lo[d], b.lo[d]); a.hi[d] = std::min(a.hi[d], b.hi[d]); }
  return a;
}
inline bool intersects(const Box& a, const Box& b) { return !(a & b).empty(); }
inline Box coarsen(Box b, int r) {
  for (int d = 0; d < 3; d++) { b.lo[d] = floordiv(b.lo[d], r); b.hi[d] = floordiv(b.hi[d], r); }
  return b;
}
inline Box refine(Box b, int r) {
  for (int d = 0; d < 3; d++) { b.lo[d] *= r; b.hi[d] = (b.hi[d] + 1) * r - 1; }
  return b;
}
// adjCellLo (side 0) / adjCellHi (side 1) with thickness n
inline Box adjCell(Box b, int d, int side, int n) {
  if (side == 0) { b.hi[d] = b.lo[d] - 1; b.lo[d] -= n; }
  else { b.lo[d] = b.hi[d] + 1; b.hi[d] += n; }
  return b;
}
inline Box shift(Box b, int d, int s) { b.lo[d] += s; b.hi[d] += s; return b; }
inline Box pointBox(const IV& p) { return Box(p, p); }

// Dense set of cells over a bounding box (DenseIntVectSet).
struct IVS {
  Box mb;
  std::vector<uint8_t> m;
  IVS() = default;
  explicit IVS(const Box& b) { define(b); }
  void define(const Box& b) {
    mb = b;
    m.assign(b.empty() ? 0 : (size_t)b.npts(), 1);
  }
  size_t off(int i, int j, int k) const {
    return ((size_t)(k - mb.lo[2]) * mb.len(1) + (j - mb.lo[1])) * mb.len(0) + (i - mb.lo[0]);
  }
  bool has(int i, int j, int k) const { return !m.empty() && mb.contains(i, j, k) && m[off(i, j, k)]; }
  bool has(const IV& p) const { return has(p[0], p[1], p[2]); }
  bool hasBox(const Box& b) const {
    for (int k = b.lo[2]; k <= b.hi[2]; k++)
      for (int j = b.lo[1]; j <= b.hi[1]; j++)
        for (int i = b.lo[0]; i <= b.hi[0]; i++)
          if (!has(i, j, k)) return false;
    return true;
  }
  template <class F> void forEach(F f) const {
    if (m.empty()) return;
    for (int k = mb.lo[2]; k <= mb.hi[2]; k++)
      for (int j = mb.lo[1]; j <= mb.hi[1]; j++)
        for (int i = mb.lo[0]; i <= mb.hi[0]; i++)
          if (m[off(i, j, k)]) f(IV{i, j, k});
  }
  void subtract(const Box& b) {
    if (m.empty()) return;
    Box c = mb & b;
    if (c.empty()) return;
    for (int k = c.lo[2]; k <= c.hi[2]; k++)
      for (int j = c.lo[1]; j <= c.hi[1]; j++)
        for (int i = c.lo[0]; i <= c.hi[0]; i++) m[off(i, j, k)] = 0;
  }
  void subtract(const IVS& o) { forEach([&](const IV& p) { if (o.has(p)) m[off(p[0], p[1], p[2])] = 0; }); }
  void intersect(const Box& b) {
    forEach([&](const IV& p) { if (!b.contains(p)) m[off(p[0], p[1], p[2])] = 0; });
  }
  long long numPts() const {
    long long c = 0;
    for (uint8_t v : m) c += v;
    return c;
  }
  Box minBox() const {
    Box r(IV{1 << 30, 1 << 30, 1 << 30}, IV{-(1 << 30), -(1 << 30), -(1 << 30)});
    bool any = false;
    forEach([&](const IV& p) {
      any = true;
      for (int d = 0; d < 3; d++) { r.lo[d] = std::min(r.lo[d], p[d]); r.hi[d] = std::max(r.hi[d], p[d]); }
    });
    return any ? r : Box();
  }
  IVS coarsened(int r) const {
    IVS c;
    if (m.empty()) return c;
    c.mb = coarsen(mb, r);
    c.m.assign((size_t)c.mb.npts(), 0);
    forEach([&](const IV& p) {
      c.m[c.off(floordiv(p[0], r), floordiv(p[1], r), floordiv(p[2], r))] = 1;
    });
    return c;
  }
  // DenseIntVectSet::grow(idir, -1)
  void erode(int d) {
    if (m.empty()) return;
    IVS o = *this;
    forEach([&](const IV& p) {
      IV a = p, b = p;
      a[d] -= 1; b[d] += 1;
      if (!(o.has(a) && o.has(b))) m[off(p[0], p[1], p[2])] = 0;
    });
  }
};

// A closed, sorted box layout (DisjointBoxLayout).
struct DBL {
  std::vector<Box> boxes;
  DBL() = default;
  explicit DBL(std::vector<Box> b) : boxes(std::move(b)) {}
  int size() const { return (int)boxes.size(); }
  const Box& operator[](int i) const { return boxes[i]; }
  long long numCells() const {
    long long n = 0;
    for (auto& b : boxes) n += b.npts();
    return n;
  }
  bool sorted() const {
    for (size_t i = 1; i < boxes.size(); i++)
      if (!(boxes[i - 1] < boxes[i])) return false;
    return true;
  }
  bool coarsenable(int r) const {
    for (auto& b : boxes)
      if (!(refine(coarsen(b, r), r) == b)) return false;
    return true;
  }
  DBL coarsened(int r) const {
    DBL c;
    c.boxes.reserve(boxes.size());
    for (auto& b : boxes) c.boxes.push_back(coarsen(b, r));
    return c;
  }
  // index of the box containing p, or -1
  int find(const IV& p) const {
    for (int q = 0; q < size(); q++)
      if (boxes[q].contains(p)) return q;
    return -1;
  }
};

}  // namespace amrmg
