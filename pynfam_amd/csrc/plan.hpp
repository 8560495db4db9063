// Host-side planner for the block-matrix products of the pnFAM iteration.
//
// The reference evaluates W_p dR W_n^T and W_p^T dH W_n as 2x2 "big block"
// products whose quadrants carry sign, transpose and imaginary flags
// (pnfam_type_bbm.f90:428-576) over block matrices with one block per block row
// (pnfam_type_blockmatrix.f90:140-211).  Here that algebra is run once,
// symbolically, on the host: the result is a list of block triple products
//   out[i,j] += s * op(A_ik) * op(B_kl) * op(C_lj)
// grouped by output block, which the device kernel executes for every
// frequency of a batch.  A and C are the resident Bogoliubov matrices; B is a
// per-frequency matrix.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pnfam {

struct BlockMap {
  std::vector<int> ir2c, ic2r, ir2m, ic2m;  // 0-based, -1 = none
  explicit BlockMap(int nb = 0) : ir2c(nb, -1), ic2r(nb, -1), ir2m(nb, -1), ic2m(nb, -1) {}
  static BlockMap from_rows(const std::vector<int> &db, const std::vector<int> &row2col) {
    BlockMap m((int)db.size());
    int off = 0;
    for (size_t i = 0; i < db.size(); ++i) {
      int c = row2col[i];
      if (c < 0) continue;
      m.ir2c[i] = c; m.ic2r[c] = (int)i; m.ir2m[i] = off; m.ic2m[c] = off;
      off += db[i] * db[c];
    }
    return m;
  }
  int64_t nelem(const std::vector<int> &db) const {
    int64_t n = 0;
    for (size_t i = 0; i < db.size(); ++i)
      if (ir2c[i] >= 0) n += (int64_t)db[i] * db[ir2c[i]];
    return n;
  }
};

// Resident matrices usable as A or C.
enum MatId { MAT_UP = 0, MAT_VP = 1, MAT_UN = 2, MAT_VN = 3, MAT_NONE = -1 };

// One device triple-product term.  Layout shared with the kernel (64 bytes).
struct Term {
  int32_t nda, ndc;            // output block rows / cols
  int32_t a_id, a_off, ta;     // A (resident), transpose flag 0/1
  int32_t b_slot, b_off, tb, ldb;
  int32_t c_id, c_off, tc;
  double s;                    // sign/alpha of the term
  int32_t pad0, pad1;
};
static_assert(sizeof(Term) == 64, "Term layout");

struct Group {                 // all terms writing one output block
  int32_t first, count;
  int32_t out_slot, out_off;
  int32_t nrow, ncol;
  int32_t pad0, pad1;
};
static_assert(sizeof(Group) == 32, "Group layout");

// Symbolic 2x2 quadrant matrix: per quadrant a map pointer (null = absent), the
// storage identity (resident id for A/C operands, per-frequency slot for B
// operands and outputs), sign and transpose flags, and the imaginary flag.
struct SymBBM {
  const BlockMap *m[4] = {nullptr, nullptr, nullptr, nullptr};
  int id[4] = {-1, -1, -1, -1};
  char t[4] = {'n', 'n', 'n', 'n'};
  double s[4] = {1, 1, 1, 1};
  bool imag = false;
};

inline SymBBM transpose_bbm(SymBBM a) {
  std::swap(a.m[1], a.m[2]);
  std::swap(a.id[1], a.id[2]);
  std::swap(a.s[1], a.s[2]);
  for (int q = 0; q < 4; ++q) a.t[q] = (a.t[q] == 'n') ? 't' : 'n';
  return a;
}

// Output quadrant of a product: map to be (re)built, and where terms go.
struct OutQuad {
  BlockMap *map;
  int slot;
};

struct Planner {
  std::vector<int> db;
  std::vector<Term> terms;      // flat, later grouped
  struct Raw { int out_slot, out_off, row; Term t; };
  std::vector<Raw> raw;
  std::string err;

  // Symbolic triprod (pnfam_type_blockmatrix.f90:140-211): rebuilds the
  // output map and records one term per output block row.
  bool triprod(char ta, const BlockMap &a, int aid, char tb, const BlockMap &b, int bslot, char tc,
               const BlockMap &c, int cid, double alpha, OutQuad out) {
    const int nb = (int)db.size();
    BlockMap nm(nb);
    int ipt = 0;
    for (int i = 0; i < nb; ++i) {
      int k, l, j, ipa, ipb, ipc;
      if (ta == 'n') { k = a.ir2c[i]; ipa = a.ir2m[i]; } else { k = a.ic2r[i]; ipa = a.ic2m[i]; }
      if (k < 0) continue;
      if (tb == 'n') { l = b.ir2c[k]; ipb = b.ir2m[k]; } else { l = b.ic2r[k]; ipb = b.ic2m[k]; }
      if (l < 0) continue;
      if (tc == 'n') { j = c.ir2c[l]; ipc = c.ir2m[l]; } else { j = c.ic2r[l]; ipc = c.ic2m[l]; }
      if (j < 0) continue;
      nm.ir2c[i] = j; nm.ic2r[j] = i; nm.ir2m[i] = ipt; nm.ic2m[j] = ipt;
      int nda = db[i], ndc = db[j];
      if (db[k] != nda || db[l] != ndc) { err = "non-square A or C block"; return false; }
      Term t{};
      t.nda = nda; t.ndc = ndc;
      t.a_id = aid; t.a_off = ipa; t.ta = ta == 't';
      t.b_slot = bslot; t.b_off = ipb; t.tb = tb == 't'; t.ldb = (tb == 'n') ? nda : ndc;
      t.c_id = cid; t.c_off = ipc; t.tc = tc == 't';
      t.s = alpha;
      raw.push_back({out.slot, ipt, i, t});
      ipt += nda * ndc;
    }
    *out.map = nm;
    return true;
  }

  // triprod_bbm_quad (pnfam_type_bbm.f90:428-552)
  bool quad(char ta, SymBBM a, char tb, SymBBM b, char tc, SymBBM c, int quadrant, double s, char tabc,
            OutQuad out) {
    if (ta == 't') a = transpose_bbm(a);
    if (tb == 't') b = transpose_bbm(b);
    if (tc == 't') c = transpose_bbm(c);
    if (tabc == 't') {
      std::swap(a, c);
      a = transpose_bbm(a); b = transpose_bbm(b); c = transpose_bbm(c);
      if (quadrant == 12) quadrant = 21; else if (quadrant == 21) quadrant = 12;
    }
    int im = (int)a.imag + (int)b.imag + (int)c.imag;
    if (im >= 2) s = -s;
    static const int T11[4][3] = {{0, 0, 0}, {1, 2, 0}, {0, 1, 2}, {1, 3, 2}};
    static const int T12[4][3] = {{0, 0, 1}, {1, 2, 1}, {0, 1, 3}, {1, 3, 3}};
    static const int T21[4][3] = {{2, 0, 0}, {3, 2, 0}, {2, 1, 2}, {3, 3, 2}};
    static const int T22[4][3] = {{2, 0, 1}, {3, 2, 1}, {2, 1, 3}, {3, 3, 3}};
    const int(*T)[3] = quadrant == 11 ? T11 : quadrant == 12 ? T12 : quadrant == 21 ? T21 : T22;
    for (int t = 0; t < 4; ++t) {
      int qa = T[t][0], qb = T[t][1], qc = T[t][2];
      if (a.m[qa] && b.m[qb] && c.m[qc])
        if (!triprod(a.t[qa], *a.m[qa], a.id[qa], b.t[qb], *b.m[qb], b.id[qb], c.t[qc], *c.m[qc], c.id[qc],
                     s * a.s[qa] * b.s[qb] * c.s[qc], out))
          return false;
    }
    return true;
  }

  // triprod_bbm (pnfam_type_bbm.f90:557-576); abc.m[q] are output maps, abc.id[q] output slots
  bool product(char ta, const SymBBM &a, char tb, const SymBBM &b, char tc, const SymBBM &c, SymBBM &abc,
               BlockMap *outmaps[4]) {
    static const int qn[4] = {11, 12, 21, 22};
    for (int q = 0; q < 4; ++q)
      if (abc.m[q])
        if (!quad(ta, a, tb, b, tc, c, qn[q], abc.s[q], abc.t[q], OutQuad{outmaps[q], abc.id[q]})) return false;
    return true;
  }

  // Group the recorded terms by output block, in output order.
  void finish(std::vector<Term> &out_terms, std::vector<Group> &groups) const {
    std::vector<Raw> r = raw;
    std::stable_sort(r.begin(), r.end(), [](const Raw &x, const Raw &y) {
      return x.out_slot != y.out_slot ? x.out_slot < y.out_slot : x.out_off < y.out_off;
    });
    out_terms.clear(); groups.clear();
    for (size_t i = 0; i < r.size();) {
      size_t j = i;
      Group g{};
      g.first = (int)out_terms.size(); g.out_slot = r[i].out_slot; g.out_off = r[i].out_off;
      g.nrow = r[i].t.nda; g.ncol = r[i].t.ndc;
      while (j < r.size() && r[j].out_slot == r[i].out_slot && r[j].out_off == r[i].out_off) {
        out_terms.push_back(r[j].t);
        ++j;
      }
      g.count = (int)(j - i);
      groups.push_back(g);
      i = j;
    }
  }
};

}  // namespace pnfam
