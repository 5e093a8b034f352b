// Adaptive mesh on the host: hierarchy construction and regrid (Mesh.cpp),
// patch connectivity (Rectangle.cpp:671-864), refinement flags
// (Rectangle.cpp:866-890, computed on the device), and the translation of the
// reference's per-rectangle ghost-fill, restriction, limiter-C, coarse-fine
// flux and regrid-transfer loops into device task programs
// (include/veritas_amd.h, vx_cell_task).
//
// A task list records the operations in the reference's order.  A new batch
// (one launch) starts whenever a task would read or write a cell written by,
// or write a cell read by, an earlier task of the current batch, so the
// parallel batches reproduce the sequential order of the reference exactly.
// The regrid transfer needs no tracking: its phases (interpolation into a new
// level, copies from the old patches of the same level, the missing cells)
// each write every cell at most once and read only cells that earlier phases
// wrote, so one batch per phase keeps the order.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iterator>
#include <limits>
#include <set>
#include <tuple>
#include <unordered_set>

#include "veritas.hpp"

#include <iomanip>
#include <sstream>

namespace vxh {

namespace {

// ---------------------------------------------------------------- task lists
enum { ARR_F = 0, ARR_CX = 1, ARR_CP = 2 };

uint64_t cell_key(int patch, int arr, int i, int j) {
    return ((uint64_t)(uint32_t)patch << 48) | ((uint64_t)arr << 44) | ((uint64_t)(uint32_t)(i + 4) << 22) |
           (uint64_t)(uint32_t)(j + 4);
}

struct TaskList {
    std::vector<vx_cell_task> tasks;
    std::vector<int32_t> starts{0};
    // Conflicts with the current batch (number `cur`): f cells of a patch in
    // dense stamp arrays over [-4, n_x+4) x [-4, n_p+4) (a cell is in the
    // batch's read or write set when its stamp equals `cur`, so a new batch
    // clears them by incrementing `cur`); limiter-C cells, and any cell outside
    // that window, in hash sets.  Without shapes every cell uses the sets.
    struct Stamps {
        std::vector<int32_t> rd, wr;
    };
    std::vector<std::pair<int, int>> shape;  // (n_x, n_p) per table index
    std::vector<Stamps> stamps;              // per table index, allocated on first use
    int32_t cur = 1;
    std::unordered_set<uint64_t> rd, wr;
    std::vector<uint64_t> r_, w_;  // hashed cells of the task being recorded

    void set_shapes(const std::vector<Rectangle *> &order) {
        shape.clear();
        for (Rectangle *x : order) shape.push_back({x->n_x, x->n_p});
        stamps.assign(shape.size(), Stamps{});
    }
    int32_t *slot(int patch, int arr, int i, int j, bool write) {
        if (arr != ARR_F || patch < 0 || patch >= (int)shape.size()) return nullptr;
        const int nx = shape[patch].first, np = shape[patch].second;
        if (i < -4 || i >= nx + 4 || j < -4 || j >= np + 4) return nullptr;
        Stamps &S = stamps[patch];
        if (S.rd.empty()) {
            S.rd.assign((size_t)(nx + 8) * (np + 8), 0);
            S.wr.assign((size_t)(nx + 8) * (np + 8), 0);
        }
        const size_t k = (size_t)(i + 4) * (np + 8) + (j + 4);
        return write ? &S.wr[k] : &S.rd[k];
    }
    void begin() {
        r_.clear();
        w_.clear();
        rd_cells_.clear();
        wr_cells_.clear();
    }
    void read(int patch, int arr, int i, int j) {
        if (slot(patch, arr, i, j, false)) rd_cells_.push_back({patch, i, j});
        else r_.push_back(cell_key(patch, arr, i, j));
    }
    void write(int patch, int arr, int i, int j) {
        if (int32_t *p = slot(patch, arr, i, j, true)) wr_cells_.push_back({patch, i, j, p});
        else w_.push_back(cell_key(patch, arr, i, j));
    }
    void commit(const vx_cell_task &t) {
        bool conflict = false;
        for (uint64_t k : r_) conflict = conflict || wr.count(k);
        for (uint64_t k : w_) conflict = conflict || wr.count(k) || rd.count(k);
        for (const Cell &c : rd_cells_) conflict = conflict || *slot(c.patch, ARR_F, c.i, c.j, true) == cur;
        for (const WCell &c : wr_cells_)
            conflict = conflict || *c.w == cur || *slot(c.patch, ARR_F, c.i, c.j, false) == cur;
        if (conflict) cut();
        tasks.push_back(t);
        rd.insert(r_.begin(), r_.end());
        wr.insert(w_.begin(), w_.end());
        for (const Cell &c : rd_cells_) *slot(c.patch, ARR_F, c.i, c.j, false) = cur;
        for (const WCell &c : wr_cells_) *c.w = cur;
        rd_cells_.clear();
        wr_cells_.clear();
    }
    void push(const vx_cell_task &t) { tasks.push_back(t); }  // untracked (the caller cuts between phases)
    void cut() {  // force a batch boundary
        if ((int32_t)tasks.size() != starts.back()) starts.push_back((int32_t)tasks.size());
        rd.clear();
        wr.clear();
        cur++;
    }
    vx_program *upload(Device &dev) {
        if (tasks.empty()) return nullptr;
        std::vector<int32_t> st = starts;
        if (st.back() != (int32_t)tasks.size()) st.push_back((int32_t)tasks.size());
        vx_program *p = nullptr;
        dev.check(vx_program_create(dev.ctx, tasks.data(), (int32_t)tasks.size(), st.data(),
                                    (int32_t)st.size() - 1, &p),
                  "program");
        return p;
    }

private:
    struct Cell {
        int patch, i, j;
    };
    struct WCell {
        int patch, i, j;
        int32_t *w;
    };
    std::vector<Cell> rd_cells_;
    std::vector<WCell> wr_cells_;
};

vx_cell_task mk(int op, int dst, int di, int dj, int src, int si, int sj, int aux) {
    vx_cell_task t;
    t.op = op;
    t.dst = dst;
    t.di = di;
    t.dj = dj;
    t.src = src;
    t.si = si;
    t.sj = sj;
    t.aux = aux;
    return t;
}

// ---------------------------------------------------------------- ghost tasks
struct GhostEmitter {
    TaskList &L;
    int r;

    // dst(di, dj) <- same-level neighbour s at level coordinates (gi, gj);
    // the sentinel gives 0 (BoundaryCondition.cpp:6-8)
    void same(Rectangle &D, int di, int dj, Rectangle *s, int gi, int gj) {
        L.begin();
        L.write(D.table_index, ARR_F, di, dj);
        if (!s) {
            L.commit(mk(VX_OP_ZERO, D.table_index, di, dj, 0, 0, 0, 0));
            return;
        }
        int si = gi - s->x_pos, sj = gj - s->p_pos;
        L.read(s->table_index, ARR_F, si, sj);
        L.commit(mk(VX_OP_COPY, D.table_index, di, dj, s->table_index, si, sj, 0));
    }
    // element e of coarse patch C's interpolation of fine cell (gi, gj)
    void coarse(Rectangle &D, int di, int dj, Rectangle &C, int gi, int gj, int e) {
        L.begin();
        L.write(D.table_index, ARR_F, di, dj);
        int ic = gi / r - C.x_pos, jc = gj / r - C.p_pos;
        for (int a = -2; a <= 2; a++)
            for (int b = -2; b <= 2; b++) L.read(C.table_index, ARR_F, ic + a, jc + b);
        L.commit(mk(VX_OP_COARSE, D.table_index, di, dj, C.table_index, gi, gj, e));
    }
    // index in the r x r interpolants of the reduced element q (2k or 2k+1)
    // for edge direction d (Rectangle.cpp:393-414)
    int full(int d, int q) const {
        int k = q / 2;
        bool second = q % 2;
        if (d == 0) return second ? r * k + r - 2 : r * k + r - 1;
        if (d == 1) return second ? r + k : k;
        if (d == 2) return second ? r * k + 1 : r * k;
        return second ? r * (r - 2) + k : r * (r - 1) + k;
    }

    // Rectangle.cpp:562-614
    void SameLevel(Rectangle &R) {
        for (unsigned i = 0; i + 2 < R.same_xm.size(); i++)
            if (R.same_xm[i + 1])
                for (int j = 0; j < r; j++) {
                    same(R, -1, i * r + j, R.ext_xm[i + 1], R.x_pos - 1, R.p_pos + i * r + j);
                    same(R, -2, i * r + j, R.ext_xm[i + 1], R.x_pos - 2, R.p_pos + i * r + j);
                }
        for (unsigned i = 0; i + 2 < R.same_xp.size(); i++)
            if (R.same_xp[i + 1])
                for (int j = 0; j < r; j++) {
                    same(R, R.n_x, i * r + j, R.ext_xp[i + 1], R.x_pos + R.n_x, R.p_pos + i * r + j);
                    same(R, R.n_x + 1, i * r + j, R.ext_xp[i + 1], R.x_pos + R.n_x + 1, R.p_pos + i * r + j);
                }
        for (unsigned i = 0; i < R.same_pm.size(); i++)
            if (R.same_pm[i])
                for (int j = 0; j < r; j++) {
                    same(R, i * r + j, -1, R.ext_pm[i], R.x_pos + i * r + j, R.p_pos - 1);
                    same(R, i * r + j, -2, R.ext_pm[i], R.x_pos + i * r + j, R.p_pos - 2);
                }
        for (unsigned i = 0; i < R.same_pp.size(); i++)
            if (R.same_pp[i])
                for (int j = 0; j < r; j++) {
                    same(R, i * r + j, R.n_p, R.ext_pp[i], R.x_pos + i * r + j, R.p_pos + R.n_p);
                    same(R, i * r + j, R.n_p + 1, R.ext_pp[i], R.x_pos + i * r + j, R.p_pos + R.n_p + 1);
                }
    }

    // the four x-edge corner blocks (Rectangle.cpp:476-559 and 1130-1214)
    void Corners(Rectangle &R) {
        const int last = (int)R.same_xm.size() - 2;
        // left edge: segment -1 covers p ghost rows -2, -1; segment `last`
        // covers n_p, n_p + 1
        for (int blk = 0; blk < 2; blk++) {
            int i = blk == 0 ? -1 : last;
            Rectangle *s = R.ext_xm[i + 1];
            int j0 = blk == 0 ? r - 2 : 0, j1 = blk == 0 ? r : 2;
            for (int j = j0; j < j1; j++) {
                if (!R.same_xm[i + 1]) {
                    coarse(R, -1, i * r + j, *s, R.x_pos - 1, R.p_pos + i * r, full(3, 2 * j));
                    coarse(R, -2, i * r + j, *s, R.x_pos - 1, R.p_pos + i * r, full(3, 2 * j + 1));
                } else {
                    same(R, -1, i * r + j, s, R.x_pos - 1, R.p_pos + i * r + j);
                    same(R, -2, i * r + j, s, R.x_pos - 2, R.p_pos + i * r + j);
                }
            }
        }
        const int lastp = (int)R.same_xp.size() - 2;
        for (int blk = 0; blk < 2; blk++) {
            int i = blk == 0 ? -1 : lastp;
            Rectangle *s = R.ext_xp[i + 1];
            int j0 = blk == 0 ? r - 2 : 0, j1 = blk == 0 ? r : 2;
            for (int j = j0; j < j1; j++) {
                if (!R.same_xp[i + 1]) {
                    coarse(R, R.n_x, i * r + j, *s, R.x_pos + R.n_x, R.p_pos + i * r, full(1, 2 * j));
                    coarse(R, R.n_x + 1, i * r + j, *s, R.x_pos + R.n_x, R.p_pos + i * r, full(1, 2 * j + 1));
                } else {
                    same(R, R.n_x, i * r + j, s, R.x_pos + R.n_x, R.p_pos + i * r + j);
                    same(R, R.n_x + 1, i * r + j, s, R.x_pos + R.n_x + 1, R.p_pos + i * r + j);
                }
            }
        }
    }

    // Rectangle.cpp:417-475 then the corner blocks
    void DifferentLevel(Rectangle &R) {
        for (unsigned i = 0; i + 2 < R.same_xm.size(); i++)
            if (!R.same_xm[i + 1])
                for (int j = 0; j < r; j++) {
                    coarse(R, -1, i * r + j, *R.ext_xm[i + 1], R.x_pos - 1, R.p_pos + i * r, full(3, 2 * j));
                    coarse(R, -2, i * r + j, *R.ext_xm[i + 1], R.x_pos - 1, R.p_pos + i * r, full(3, 2 * j + 1));
                }
        for (unsigned i = 0; i + 2 < R.same_xp.size(); i++)
            if (!R.same_xp[i + 1])
                for (int j = 0; j < r; j++) {
                    coarse(R, R.n_x, i * r + j, *R.ext_xp[i + 1], R.x_pos + R.n_x, R.p_pos + i * r, full(1, 2 * j));
                    coarse(R, R.n_x + 1, i * r + j, *R.ext_xp[i + 1], R.x_pos + R.n_x, R.p_pos + i * r,
                           full(1, 2 * j + 1));
                }
        for (unsigned i = 0; i < R.same_pm.size(); i++)
            if (!R.same_pm[i])
                for (int j = 0; j < r; j++) {
                    coarse(R, i * r + j, -1, *R.ext_pm[i], R.x_pos + i * r, R.p_pos - 1, full(0, 2 * j));
                    coarse(R, i * r + j, -2, *R.ext_pm[i], R.x_pos + i * r, R.p_pos - 1, full(0, 2 * j + 1));
                }
        for (unsigned i = 0; i < R.same_pp.size(); i++)
            if (!R.same_pp[i])
                for (int j = 0; j < r; j++) {
                    coarse(R, i * r + j, R.n_p, *R.ext_pp[i], R.x_pos + i * r, R.p_pos + R.n_p, full(2, 2 * j));
                    coarse(R, i * r + j, R.n_p + 1, *R.ext_pp[i], R.x_pos + i * r, R.p_pos + R.n_p,
                           full(2, 2 * j + 1));
                }
        Corners(R);
    }

    // Rectangle.cpp:329-337
    void Interior(Rectangle &R) {
        for (int i = 0; i < R.n_x; i++)
            for (int j = 0; j < R.n_p; j++)
                if (R.is_nested[R.NS(i, j)]) {
                    Rectangle &F = *R.finer_level[R.NS(i, j)];
                    int gi = R.x_pos + i, gj = R.p_pos + j;
                    L.begin();
                    L.write(R.table_index, ARR_F, i, j);
                    int i0 = gi * r - F.x_pos, j0 = gj * r - F.p_pos;
                    for (int k = 0; k < r; k++)
                        for (int l = 0; l < r; l++) L.read(F.table_index, ARR_F, i0 + k, j0 + l);
                    L.commit(mk(VX_OP_RESTRICT, R.table_index, i, j, F.table_index, gi, gj, 0));
                }
    }
};

// ---------------------------------------------------------------- limiter-C tasks
struct CEmitter {
    TaskList &L;
    int r;

    void face_pair(int op, Rectangle &N, int di, int dj, Rectangle &Q, int si, int sj, int t) {
        int arr = t < 2 ? ARR_CX : ARR_CP;
        L.begin();
        if (op == VX_OP_C_UPDATE) {
            L.read(N.table_index, arr, di, dj);
            L.read(Q.table_index, arr, si, sj);
        }
        L.write(N.table_index, arr, di, dj);
        L.write(Q.table_index, arr, si, sj);
        L.commit(mk(op, N.table_index, di, dj, Q.table_index, si, sj, t));
    }
    // SetCFromSameLevel / UpdateCFromSameLevel called on neighbour N by
    // caller Q at level coordinates (gi, gj) (Rectangle.cpp:1794-1830, 1879-1915)
    void same(int op, Rectangle &N, Rectangle &Q, int gi, int gj, int t) {
        int in = gi - N.x_pos, jn = gj - N.p_pos, ic = gi - Q.x_pos, jc = gj - Q.p_pos;
        for (int k = 0; k < r; k++) {
            if (t < 2) face_pair(op, N, in, jn + k, Q, ic, jc + k, t);
            else face_pair(op, N, in + k, jn, Q, ic + k, jc, t);
        }
    }
    // Rectangle.cpp:1832-1877 / 1917-1960; a sentinel neighbour is skipped (the
    // reference's call on it is undefined behaviour and touches wall faces only)
    void SameBoundary(int op, Rectangle &R) {
        for (unsigned i = 0; i + 2 < R.same_xm.size(); i++)
            if (R.same_xm[i + 1] && R.ext_xm[i + 1]) same(op, *R.ext_xm[i + 1], R, R.x_pos, R.p_pos + i * r, 0);
        for (unsigned i = 0; i + 2 < R.same_xp.size(); i++)
            if (R.same_xp[i + 1] && R.ext_xp[i + 1])
                same(op, *R.ext_xp[i + 1], R, R.x_pos + R.n_x, R.p_pos + i * r, 1);
        for (unsigned i = 0; i < R.same_pm.size(); i++)
            if (R.same_pm[i] && R.ext_pm[i]) same(op, *R.ext_pm[i], R, R.x_pos + i * r, R.p_pos, 2);
        for (unsigned i = 0; i < R.same_pp.size(); i++)
            if (R.same_pp[i] && R.ext_pp[i]) same(op, *R.ext_pp[i], R, R.x_pos + i * r, R.p_pos + R.n_p, 3);
    }
    // SetCFromDifferentLevel on coarse K by finer caller Q (Rectangle.cpp:1625-1706)
    void diff(Rectangle &K, Rectangle &Q, int gi, int gj, int t) {
        int icl = gi - Q.x_pos, jcl = gj - Q.p_pos;
        int ik = gi / r - K.x_pos, jk = gj / r - K.p_pos;
        int arr = t < 2 ? ARR_CX : ARR_CP;
        L.begin();
        L.read(K.table_index, arr, ik, jk);
        L.write(K.table_index, arr, ik, jk);
        for (int k = 0; k < r; k++) {
            if (t < 2) L.write(Q.table_index, arr, icl, jcl + k);
            else L.write(Q.table_index, arr, icl + k, jcl);
        }
        L.commit(mk(VX_OP_C_DIFF, K.table_index, ik, jk, Q.table_index, icl, jcl, t));
    }
    // Rectangle.cpp:1708-1760
    void DifferentBoundary(Rectangle &R) {
        for (unsigned i = 0; i + 2 < R.same_xm.size(); i++)
            if (!R.same_xm[i + 1]) diff(*R.ext_xm[i + 1], R, R.x_pos, R.p_pos + r * i, 0);
        for (unsigned i = 0; i + 2 < R.same_xp.size(); i++)
            if (!R.same_xp[i + 1]) diff(*R.ext_xp[i + 1], R, R.x_pos + R.n_x, R.p_pos + r * i, 1);
        for (unsigned i = 0; i < R.same_pm.size(); i++)
            if (!R.same_pm[i]) diff(*R.ext_pm[i], R, R.x_pos + r * i, R.p_pos, 2);
        for (unsigned i = 0; i < R.same_pp.size(); i++)
            if (!R.same_pp[i]) diff(*R.ext_pp[i], R, R.x_pos + r * i, R.p_pos + R.n_p, 3);
    }
};

// ---------------------------------------------------------------- clustering
// Berger-Rigoutsos-style recursive bisection of flagged cells (Mesh.cpp:361-792)
int cell_count(const rect &b) {
    return (b.second.first - b.first.first + 1) * (b.second.second - b.first.second + 1);
}

struct Signatures {
    std::vector<int> sx, dsx, sp, dsp;  // flag counts per column / row and their Laplacians
};

Signatures signatures(const rect &b, const std::vector<coords> &fl) {
    Signatures s;
    double lx = b.second.first - b.first.first, lp = b.second.second - b.first.second;
    s.sx.assign((size_t)(lx + 1), 0);
    s.sp.assign((size_t)(lp + 1), 0);
    for (const coords &c : fl) {
        s.sx.at(c.first - b.first.first)++;
        s.sp.at(c.second - b.first.second)++;
    }
    auto lap = [](const std::vector<int> &v, double len, std::vector<int> &out) {
        if (len >= 2) {
            out.resize((size_t)(len - 1));
            for (unsigned h = 1; h + 1 < v.size(); h++) out.at(h - 1) = v.at(h - 1) + v.at(h + 1) - 2 * v.at(h);
        }
    };
    lap(s.sx, lx, s.dsx);
    lap(s.sp, lp, s.dsp);
    return s;
}

int sign_of(int v) { return (v > 0) ? 1 : ((v < 0) ? -1 : 0); }

// strongest sign change of one Laplacian signature (Mesh.cpp:460-528 / 530-580)
struct Inflection {
    int index = 0, size = 0;
    bool bisect = false;
};
Inflection strongest(const std::vector<int> &d, int far) {
    Inflection out;
    if (d.size() <= 2) return out;
    std::vector<int> idx, mag;
    for (unsigned i = 0; i + 1 < d.size(); i++)
        if (d[i] != 0 && d[i + 1] != 0 && sign_of(d[i]) != sign_of(d[i + 1])) {
            idx.push_back((int)i);
            mag.push_back(std::abs(d[i + 1] - d[i]));
        }
    if (idx.size() == 1) {
        out.index = idx[0];
        out.size = mag[0];
    } else if (idx.size() > 1) {
        auto top = std::max_element(mag.begin(), mag.end());
        if (std::count(mag.begin(), mag.end(), *top) > 1) {  // ties: the one nearest the centre
            int centre = (int)d.size() / 2, best = far, pick = 0;
            for (unsigned i = 0; i < idx.size(); i++)
                if (best > std::abs(idx[i] - centre)) {
                    best = std::abs(idx[i] - centre);
                    pick = (int)i;
                }
            out.index = idx[pick];
            out.size = mag[pick];
        } else {
            out.index = idx[top - mag.begin()];
            out.size = mag[top - mag.begin()];
        }
    } else {
        out.bisect = true;
    }
    return out;
}

// (0, i): cut in x after column i; (1, j): cut in p; (2, 0): bisect; (3, 0): keep
coords inflection_cut(const rect &b, const Signatures &s) {
    Inflection ix = strongest(s.dsx, std::numeric_limits<int>::max());
    Inflection ip = strongest(s.dsp, 32767);
    if (ix.bisect && ip.bisect) return {2, 0};
    if ((s.dsx.size() <= 2 && s.dsp.size() <= 2) || (ix.size == 0 && ip.size == 0)) return {3, 0};
    if (ix.size > ip.size) return {0, ix.index + b.first.first + 1};
    return {1, ip.index + b.first.second + 1};
}

// first run of zeros in a signature (Mesh.cpp:597-624)
struct Hole {
    bool found = false;
    int start = 0, finish = 0;
};
Hole first_hole(const std::vector<int> &sig) {
    Hole h;
    auto z = std::find(sig.begin(), sig.end(), 0);
    if (z == sig.end()) return h;
    h.found = true;
    h.start = (int)(z - sig.begin()) - (z == sig.begin() ? 0 : 1);
    auto nz = std::find_if(z, sig.end(), [](int v) { return v != 0; });
    h.finish = nz == sig.end() ? (int)(nz - sig.begin()) - 1 : (int)(nz - sig.begin());
    return h;
}

rect bounding_box(const std::vector<coords> &fl) {  // Mesh.cpp:626-642
    int x0 = fl[0].first, x1 = fl[0].first, p0 = fl[0].second, p1 = fl[0].second;
    for (const coords &c : fl) {
        x0 = std::min(x0, c.first);
        x1 = std::max(x1, c.first);
        p0 = std::min(p0, c.second);
        p1 = std::max(p1, c.second);
    }
    return {{x0, p0}, {x1, p1}};
}

level split_box(const rect &b, const std::vector<coords> &fl, double minEfficiency) {  // Mesh.cpp:644-792
    if (fl.size() / double(cell_count(b)) > minEfficiency) return {b};
    Signatures s = signatures(b, fl);
    Hole hx = first_hole(s.sx), hp = first_hole(s.sp);
    hx.start += b.first.first;
    hx.finish += b.first.first;
    hp.start += b.first.second;
    hp.finish += b.first.second;
    rect one, two;
    auto cut_x_hole = [&]() {
        one = {b.first, {hx.start, b.second.second}};
        two = {{hx.finish, b.first.second}, b.second};
    };
    auto cut_p_hole = [&]() {
        one = {b.first, {b.second.first, hp.start}};
        two = {{b.first.first, hp.finish}, b.second};
    };
    if (hx.found && hp.found) {
        if (std::abs(hx.finish - hx.start) > std::abs(hp.finish - hp.start)) cut_x_hole();
        else cut_p_hole();
    } else if (hx.found) {
        cut_x_hole();
    } else if (hp.found) {
        cut_p_hole();
    } else {
        coords c = inflection_cut(b, s);
        if (c.first == 0) {
            one = {b.first, {c.second, b.second.second}};
            two = {{c.second + 1, b.first.second}, b.second};
        } else if (c.first == 1) {
            one = {b.first, {b.second.first, c.second}};
            two = {{b.first.first, c.second + 1}, b.second};
        } else if (c.first == 2) {
            double lx = b.second.first - b.first.first, lp = b.second.second - b.first.second;
            if (lx > lp) {
                double cut = std::floor((lx / 2) + b.first.first);
                one = {b.first, {(int)cut, b.second.second}};
                two = {{(int)(cut + 1), b.first.second}, b.second};
            } else {
                double cut = std::floor((lp / 2) + b.first.second);
                one = {b.first, {b.second.first, (int)cut}};
                two = {{b.first.first, (int)(cut + 1)}, b.second};
            }
        } else {
            return {b};
        }
    }
    std::vector<coords> f1, f2;
    for (const coords &c : fl) {
        bool in = c.first >= one.first.first && c.first <= one.second.first && c.second >= one.first.second &&
                  c.second <= one.second.second;
        (in ? f1 : f2).push_back(c);
    }
    level out, b2;
    if (!f1.empty()) out = split_box(one, f1, minEfficiency);
    if (!f2.empty()) b2 = split_box(two, f2, minEfficiency);
    out.insert(out.end(), b2.begin(), b2.end());
    return out;
}

}  // namespace

std::vector<std::vector<int32_t>> Mesh::ProgramBatches() const {
    auto sizes = [](const vx_program *p) {
        std::vector<int32_t> v(vx_program_batch_sizes(p, nullptr, 0));
        vx_program_batch_sizes(p, v.data(), (int32_t)v.size());
        return v;
    };
    std::vector<std::vector<int32_t>> out = {sizes(push_prog), sizes(pushc_prog)};
    for (auto &l : levels) out.push_back(sizes(l->flux_prog));
    return out;
}

level SplitBox(const rect &b, const std::vector<coords> &flagged, double minEfficiency) {
    return split_box(b, flagged, minEfficiency);
}

// ---------------------------------------------------------------- Rectangle links

// Rectangle.cpp:667-669
static bool on_line(int x, int y1, int y2) { return ((x - y1) > -1) && ((y2 - x) > -1); }

// Rectangle.cpp:671-766: a same-level neighbour across each edge
void Rectangle::CalculateConnectivitySame(Rectangle *o) {
    const int r = settings.refinementRatio;
    const int dx = o->x_pos - x_pos, dp = o->p_pos - p_pos;
    auto link = [&](std::vector<Rectangle *> &ext, std::vector<char> &same, size_t k) {
        ext.at(k) = o;
        same.at(k) = 1;
    };
    if (dx == n_x || dx == -o->n_x) {  // neighbour to the right or left
        auto &ext = dx == n_x ? ext_xp : ext_xm;
        auto &same = dx == n_x ? same_xp : same_xm;
        int lo = std::max(o->p_pos, p_pos), hi = std::min(p_pos + n_p, o->p_pos + o->n_p);
        if (on_line(-1, dp, dp + o->n_p - 1)) link(ext, same, 0);
        if (on_line(n_p, dp, dp + o->n_p - 1)) link(ext, same, ext.size() - 1);
        for (; lo < hi; lo += r) link(ext, same, (lo - p_pos) / r + 1);
    }
    if (dp == n_p || dp == -o->n_p) {  // neighbour above or below
        auto &ext = dp == n_p ? ext_pp : ext_pm;
        auto &same = dp == n_p ? same_pp : same_pm;
        int lo = std::max(o->x_pos, x_pos), hi = std::min(x_pos + n_x, o->x_pos + o->n_x);
        // the x-edge corner entries: last entries for a neighbour above, first below
        size_t corner = dp == n_p ? ext_xm.size() - 1 : 0;
        if (on_line(-1, dx, dx + o->n_x - 1)) link(ext_xm, same_xm, corner);
        if (on_line(n_x, dx, dx + o->n_x - 1)) link(ext_xp, same_xp, dp == n_p ? ext_xp.size() - 1 : 0);
        for (; lo < hi; lo += r) link(ext, same, (lo - x_pos) / r);
    }
}

// Rectangle.cpp:768-864: this coarse patch covers part of finer patch f
void Rectangle::CalculateConnectivityFromFiner(Rectangle *current, Rectangle *f) {
    const int r = settings.refinementRatio;
    const int fx = f->x_pos / r - x_pos, fp = f->p_pos / r - p_pos;
    const int fnx = f->n_x / r, fnp = f->n_p / r;
    const int lx = std::max(0, fx), ux = std::min(n_x, fx + fnx);
    const int lp = std::max(0, fp), up_ = std::min(n_p, fp + fnp);
    for (int i = lx; i < ux; i++)
        for (int j = lp; j < up_; j++) {
            is_nested[NS(i, j)] = 1;
            finer_level[NS(i, j)] = f;
        }
    auto mark_x = [&](int face) {  // coarse x-faces on the finer patch's edge
        for (int j = lp; j < up_; j++) {
            finer_level_x[NS(face, j)] = f;
            is_ilb_x[NS(face, j)] = 1;
        }
    };
    auto link_x = [&](std::vector<Rectangle *> &ext, std::vector<char> &same) {
        for (int j = lp; j < up_; j++) {
            ext[j - fp + 1] = current;
            same[j - fp + 1] = 0;
        }
        if (on_line(fp - 1, 0, n_p - 1)) {
            ext[0] = current;
            same[0] = 0;
        }
        if (on_line(fp + fnp, 0, n_p - 1)) {
            ext[ext.size() - 1] = current;
            same[ext.size() - 1] = 0;
        }
    };
    if (fx + fnx < n_x + 1 && fx + fnx > -1) mark_x(fx + fnx);
    if (fx + fnx < n_x && fx + fnx > -1) link_x(f->ext_xp, f->same_xp);
    if (fx < n_x + 1 && fx > -1) mark_x(fx);
    if (fx < n_x + 1 && fx > 0) link_x(f->ext_xm, f->same_xm);
    auto mark_p = [&](int face) {
        for (int i = lx; i < ux; i++) {
            finer_level_p[NS(i, face)] = f;
            is_ilb_p[NS(i, face)] = 1;
        }
    };
    auto link_p = [&](std::vector<Rectangle *> &ext, std::vector<char> &same) {
        for (int i = lx; i < ux; i++) {
            ext[i - fx] = current;
            same[i - fx] = 0;
        }
    };
    if (fp + fnp < n_p + 1 && fp + fnp > -1) mark_p(fp + fnp);
    if (fp + fnp < n_p && fp + fnp > -1) link_p(f->ext_pp, f->same_pp);
    if (fp < n_p + 1 && fp > -1) mark_p(fp);
    if (fp < n_p + 1 && fp > 0) link_p(f->ext_pm, f->same_pm);
}

// Rectangle.cpp:866-890: the error estimate runs on the device over the
// patch's inset range; flags come back in (i, j) order
void Rectangle::getError(std::vector<coords> &flagged) {
    int offset = 0;
    for (int i = 0; i <= settings.maxDepth - depth; i++) offset += std::pow(settings.refinementRatio, i);
    offset *= 3;
    const int lx = std::max(x_pos, offset), ux = std::min(x_pos + n_x, settings.GetXSize(depth) - offset);
    const int lp = std::max(p_pos, offset), up_ = std::min(p_pos + n_p, settings.GetPSize(depth, particleType) - offset);
    if (ux <= lx || up_ <= lp) return;
    const size_t n = (size_t)(ux - lx) * (up_ - lp);
    const bool hook = settings.HasRefinementOverride();
    if (!hook) {
        // device compaction: only the flagged cells (and the range's signatures)
        // come back, in the row-major order of the scan below
        std::vector<int32_t> cells(n);
        int32_t count = 0;
        sig_x.assign(ux - lx, 0);
        sig_p.assign(up_ - lp, 0);
        dev->check(vx_flag_compact(dev->ctx, &P, errorWeights, settings.refinementCriteria, lx - x_pos, ux - x_pos,
                                   lp - p_pos, up_ - p_pos, cells.data(), (int32_t)n, &count, sig_x.data(),
                                   sig_p.data()),
                   "flag compaction");
        const int wd = up_ - lp;
        for (int32_t k = 0; k < count; k++) flagged.push_back(std::make_pair(lx + cells[k] / wd, lp + cells[k] % wd));
        sig_lo_x = lx;
        sig_lo_p = lp;
        return;
    }
    void *d = nullptr;
    dev->check(vx_malloc(dev->ctx, n, &d), "alloc flags");
    std::vector<uint8_t> h(n);
    int rc = vx_flag_cells(dev->ctx, &P, errorWeights, settings.refinementCriteria, lx - x_pos, ux - x_pos,
                           lp - p_pos, up_ - p_pos, (uint8_t *)d);
    if (rc == VX_OK) rc = vx_memcpy_d2h(dev->ctx, h.data(), d, n);
    vx_free(dev->ctx, d);
    dev->check(rc, "flags");
    size_t e = 0;
    for (int i = lx; i < ux; i++)
        for (int j = lp; j < up_; j++, e++)
            if (h[e] || (hook && settings.RefinementOverride((i + 0.5) * P.dx, P.pmin + P.dp * (j - p_pos + p_pos),
                                                                  depth, particleType)))
                flagged.push_back(std::make_pair(i, j));
}

void Rectangle::UploadNested() {
    bool any = std::find(is_nested.begin(), is_nested.end(), 1) != is_nested.end();
    if (!any) {
        P.nested = nullptr;
        return;
    }
    const int pitch = P.pitch;
    std::vector<uint8_t> h(vx_plane_elems(n_x, n_p), 0);
    for (int i = 0; i < n_x; i++)
        for (int j = 0; j < n_p; j++) h[(size_t)(i + 2) * pitch + (j + 2)] = is_nested[NS(i, j)] ? 1 : 0;
    if (!nested_d) {
        void *d = nullptr;
        dev->check(vx_malloc(dev->ctx, h.size(), &d), "alloc nested");
        allocations.push_back(d);
        nested_d = (uint8_t *)d;
    }
    dev->check(vx_memcpy_h2d(dev->ctx, nested_d, h.data(), h.size()), "upload nested");
    P.nested = nested_d;
}

// ---------------------------------------------------------------- Level programs
void Level::FreePrograms() {
    for (int k = 0; k < 7; k++) {
        vx_program_destroy(prog[k]);
        prog[k] = nullptr;
        prog_built[k] = false;
    }
    vx_program_destroy(flux_prog);
    flux_prog = nullptr;
}

Level::~Level() { FreePrograms(); }

// tasks of update type `type` (Level.cpp:88-126) of every patch of level L, appended to T
static void emit_level(TaskList &T, Level &L, int type, int r) {
    GhostEmitter g{T, r};
    CEmitter c{T, r};
    for (auto &R : L.rectangles) {
        switch (type) {
            case 0: g.Interior(*R); break;
            case 1: g.SameLevel(*R); break;
            case 2: g.DifferentLevel(*R); break;
            case 3: g.Corners(*R); break;
            case 4: c.SameBoundary(VX_OP_C_SAME, *R); break;
            case 5: c.DifferentBoundary(*R); break;
            case 6: c.SameBoundary(VX_OP_C_UPDATE, *R); break;
        }
    }
}

// Level.cpp:88-126 (the mesh runs its merged programs; a direct call builds
// this level's program of the type on first use)
void Level::PushData(int updateType, int val) {
    if (updateType < 0 || updateType > 6) throw Error("Level::PushData: bad update type");
    if (!table || !dev) throw Error("Level::PushData: level not attached to a mesh");
    if (!prog_built[updateType]) {
        TaskList T;
        emit_level(T, *this, updateType, settings.refinementRatio);
        prog[updateType] = T.upload(*dev);
        prog_built[updateType] = true;
    }
    vx_program *p = prog[updateType];
    if (!p) return;  // nothing to do on this level
    if (updateType <= 3)
        dev->check(vx_program_run(dev->ctx, table, p, VX_MODE_GHOST, val, settings.refinementRatio), "push data");
    else
        dev->check(vx_program_run(dev->ctx, table, p, VX_MODE_GHOST, 1, settings.refinementRatio), "limiter sync");
}

void Level::RunFluxProgram(int step, double dt, const vx_fields &fl) {
    if (!flux_prog) return;
    dev->check(vx_program_run_fluxes(dev->ctx, table, flux_prog, step, dt, settings.refinementRatio, &fl),
               "coarse-fine fluxes");
}

// ---------------------------------------------------------------- Mesh

// Mesh.cpp:8-50.  With MKLINIT == 0 (veritas.hpp:33) the initial error pass
// flags nothing, so the hierarchy grows through updateHierarchy(true).
Mesh::Mesh(int sp, Settings &s, std::shared_ptr<Device> d) : particleType(sp), settings(s), dev(std::move(d)) {
    int xmax = s.GetXSize(s.maxDepth), pmax = s.GetPSize(s.maxDepth, sp);
    hierarchy.push_back(level{rect{{0, 0}, {xmax, pmax}}});
    promoteHierarchyToMesh(true);
    for (int i = 0; i < s.maxDepth; i++) {
        PushData();
        updateHierarchy(true);
    }
}

Mesh::~Mesh() {
    for (auto &l : levels) l->FreePrograms();
    vx_patch_table_destroy(table);
}

bool Mesh::SinglePatch() const {
    int n = 0;
    for (auto &l : levels) n += (int)l->rectangles.size();
    return n == 1;
}

// Mesh.cpp:212-295 (time-advance branch; the initial branch flags nothing)
void Mesh::getError(int lvl, bool init, std::vector<coords> &flagged) {
    if (init) return;
    for (auto &r : levels.at(settings.maxDepth - lvl)->rectangles) r->getError(flagged);
}

namespace {
// Mesh.cpp:298-313: boxes found on level lvl, as exclusive ranges on level lvl+1
void boxes_up(Settings &s, int sp, level &boxes, int lvl) {
    double rx = s.GetXSize(s.maxDepth - (lvl + 1)) / double(s.GetXSize(s.maxDepth - lvl));
    double rp = s.GetPSize(s.maxDepth - (lvl + 1), sp) / double(s.GetPSize(s.maxDepth - lvl, sp));
    for (rect &b : boxes) {
        b.second.first += 1;
        b.second.second += 1;
        b.first.first *= rx;
        b.first.second *= rp;
        b.second.first *= rx;
        b.second.second *= rp;
    }
}

}  // namespace

// Mesh.cpp:315-339: level-lvl cells under a level-(lvl+2) box grown by 2r, sorted
// and unique, for the box [i0, i1] x [j0, j1] of the finer level (i1, j1
// inclusive, already grown) on a level of nmax x pmax cells, fx = nmax /
// n(lvl+2), fp likewise.  The reference visits every level-(lvl+2) cell and maps
// it with trunc(i * fx); fx < 1 and the map is monotone with steps of at most
// 1, so its image over [a, b] is every integer from trunc(a * fx) to
// trunc(b * fx): the same list is the rectangle of those ranges (clipped to the
// level), written in sorted order without visiting the fine cells.
void cells_under_box(int nmax, int pmax, double fx, double fp, int i0, int i1, int j0, int j1,
                     std::vector<coords> &out) {
    if (fx > 1.0 || fp > 1.0 || i1 < i0 || j1 < j0) {  // not a coarser level: the reference's loop
        for (int i = i0; i <= i1; i++)
            for (int j = j0; j <= j1; j++) {
                int ic = i * fx, jc = j * fp;
                if (ic > -1 && ic < nmax && jc > -1 && jc < pmax) out.push_back({ic, jc});
            }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
        return;
    }
    const int ic0 = std::max(0, (int)(i0 * fx)), ic1 = std::min(nmax - 1, (int)(i1 * fx));
    const int jc0 = std::max(0, (int)(j0 * fp)), jc1 = std::min(pmax - 1, (int)(j1 * fp));
    for (int ic = ic0; ic <= ic1; ic++)
        for (int jc = jc0; jc <= jc1; jc++) out.push_back({ic, jc});
}

namespace {
void cells_under(Settings &s, int sp, int lvl, const rect &b, std::vector<coords> &out) {
    const int nmax = s.GetXSize(s.maxDepth - lvl), pmax = s.GetPSize(s.maxDepth - lvl, sp);
    const double fx = nmax / double(s.GetXSize(s.maxDepth - (lvl + 2)));
    const double fp = pmax / double(s.GetPSize(s.maxDepth - (lvl + 2), sp));
    const int R = (int)s.refinementRatio;
    cells_under_box(nmax, pmax, fx, fp, b.first.first - 2 * R, b.second.first + 2 * R + 1, b.first.second - 2 * R,
                    b.second.second + 2 * R, out);
}
}  // namespace

// Mesh.cpp:132-210.  When the new hierarchy equals the current one, the
// reference's rebuild would reproduce the current state: every patch takes its
// interior back from the old patch of the same box, the ghost cells are refilled
// by the same PushData from the same data, f^n is copied from f again, and
// every other plane is rewritten before it is read in the next stage.  The
// patches are then kept as they are.
void Mesh::updateHierarchy(bool init) {
    if (hierarchy.empty()) throw Error("hierarchy is empty, cannot update it");
    const std::vector<level> before = hierarchy;
    const int top = (int)hierarchy.size() - 1;
    for (int l = top; l > -1; l--) {
        // at the deepest allowed level the reference flags and clusters, then
        // discards the boxes: nothing to compute
        if (l == top && l >= settings.maxDepth) continue;
        std::vector<coords> flagged;
        if (l < (int)hierarchy.size() - 2) {  // keep level l+2 nested inside the new level l+1
            std::vector<coords> own, below;
            if (hierarchy.size() > 2)
                for (const rect &b : hierarchy.at(l + 2)) {
                    std::vector<coords> t;
                    cells_under(settings, particleType, l, b, t);
                    below.insert(below.end(), t.begin(), t.end());
                }
            getError(l, false, own);
            std::sort(own.begin(), own.end());
            std::set_union(own.begin(), own.end(), below.begin(), below.end(), std::back_inserter(flagged));
        } else {
            getError(l, false, flagged);
        }
        if (!flagged.empty()) {
            level boxes = split_box(bounding_box(flagged), flagged, settings.minEfficiency);
            if (l == top) {
                if (l < settings.maxDepth) {
                    boxes_up(settings, particleType, boxes, l);
                    hierarchy.push_back(boxes);
                }
            } else {
                boxes_up(settings, particleType, boxes, l);
                hierarchy.at(l + 1) = boxes;
            }
        } else if (l < top) {
            hierarchy.erase(hierarchy.begin() + l + 1);
        }
    }
    if (!init) regrids++;
    if (!init && hierarchy == before && !levels.empty()) return;
    if (!init) rebuilds++;
    promoteHierarchyToMesh(init);
}

// Mesh.cpp:794-875.  A level's extent is taken at N - lvl with N + 1 the
// number of levels in use, so while levels are unused the up/right walls are
// compared with a finer level's size (kept as the reference has it).
void Mesh::promoteHierarchyToMesh(bool init) {
    const int N = (int)hierarchy.size() - 1;
    const int empty = settings.maxDepth - N;
    std::vector<std::unique_ptr<Level>> old;
    if (!init)
        for (auto &l : levels) old.push_back(std::move(l));
    for (auto &l : old) l->FreePrograms();
    levels.clear();
    for (int i = 0; i < empty; i++) levels.push_back(std::unique_ptr<Level>(new Level(particleType, i, settings)));
    for (int lvl = N; lvl >= 0; lvl--) {
        const int xmax = settings.GetXSize(N - lvl), pmax = settings.GetPSize(N - lvl, particleType);
        const int depth = N - lvl + empty;
        levels.push_back(std::unique_ptr<Level>(new Level(particleType, depth, settings)));
        for (const rect &b : hierarchy.at(lvl)) {
            bool up = b.second.second == pmax, down = b.first.second == 0;
            bool left = b.first.first == 0, right = b.second.first == xmax;
            levels.at(depth)->rectangles.push_back(std::make_shared<Rectangle>(
                b.second.first - b.first.first, b.second.second - b.first.second, b.first.first, b.first.second,
                depth, settings, dev, up, down, left, right, particleType));
        }
    }
    if (init)
        for (auto &l : levels)
            for (auto &r : l->rectangles) r->InitializeDistribution();
    for (unsigned l = 1; l < levels.size(); l++)
        for (auto &rc : levels[l]->rectangles)
            for (auto &rf : levels[l - 1]->rectangles) rc->CalculateConnectivityFromFiner(rc.get(), rf.get());
    for (auto &l : levels)
        for (auto &a : l->rectangles)
            for (auto &b : l->rectangles)
                if (a != b) a->CalculateConnectivitySame(b.get());
    for (auto &l : levels)
        for (auto &r : l->rectangles) r->UploadNested();
    SetFieldSolver(EMSolver);
    BuildPrograms();
    if (!init) InterMeshDataTransfer(old);
    old.clear();
    PushData();
    for (auto &l : levels)
        for (auto &r : l->rectangles) r->FCTTimeStep(0, -1, 3);
}

void Mesh::FreePrograms() {
    for (auto &l : levels) l->FreePrograms();
    vx_program_destroy(push_prog);
    vx_program_destroy(pushc_prog);
    vx_program_destroy(pushcc_prog);
    push_prog = pushc_prog = pushcc_prog = nullptr;
    vx_patch_table_destroy(table);
    table = nullptr;
}

void Mesh::RefreshTable() {
    std::vector<vx_patch> d;
    for (Rectangle *r : order) d.push_back(r->desc());
    if (table) dev->check(vx_patch_table_update(dev->ctx, table, d.data(), (int32_t)d.size()), "patch table");
}

// Patch table of this mesh (levels finest first) and one program per level
// and update type, plus the coarse-fine flux program of each coarse level.
void Mesh::BuildPrograms() {
    FreePrograms();
    order.clear();
    for (auto &l : levels)
        for (auto &r : l->rectangles) {
            r->UnaliasSlots();
            r->table_index = (int)order.size();
            order.push_back(r.get());
        }
    std::vector<vx_patch> d;
    for (Rectangle *r : order) d.push_back(r->desc());
    dev->check(vx_patch_table_create(dev->ctx, d.data(), (int32_t)d.size(), &table), "patch table");
    const int r = settings.refinementRatio;
    // Mesh::PushData's order of level programs (Mesh.cpp:91-106) as one task
    // list, and PushBoundaryC's (Mesh.cpp:904-917): a task joins the current
    // batch only when it touches no cell that the batch reads or writes, so
    // the merged program applies every task in the sequential order's effect
    {
        TaskList P;
        P.set_shapes(order);
        emit_level(P, *levels[0], 1, r);
        for (unsigned i = 1; i < levels.size(); i++) {
            emit_level(P, *levels[i], 0, r);
            emit_level(P, *levels[i], 1, r);
        }
        emit_level(P, *levels.back(), 3, r);
        for (unsigned i = levels.size() - 1; i > 0; i--) emit_level(P, *levels[i - 1], 2, r);
        push_prog = P.upload(*dev);
        TaskList C;
        C.set_shapes(order);
        for (int type = 4; type <= 6; type++)
            for (auto &l : levels) emit_level(C, *l, type, r);
        pushc_prog = C.upload(*dev);
        // the same, then the correction of every interior cell whose
        // corrected value reads a C that the synchronisation writes: Cx at
        // (i, j) is read by cells (i, j) and (i - 1, j), Cp at (i, j) by (i, j)
        // and (i, j - 1) (vxi::corrected_cell); in patch, row and column order
        std::set<std::tuple<int, int, int>> fix;
        auto face = [&](int p, bool x, int i, int j) {
            const Rectangle &R = *order.at(p);
            for (int k = 0; k < 2; k++) {
                const int ci = x ? i - k : i, cj = x ? j : j - k;
                if (ci >= 0 && ci < R.n_x && cj >= 0 && cj < R.n_p) fix.insert({p, ci, cj});
            }
        };
        for (const vx_cell_task &t : C.tasks) {
            const bool x = t.aux < 2;
            if (t.op == VX_OP_C_SAME || t.op == VX_OP_C_UPDATE) {
                face(t.dst, x, t.di, t.dj);
                face(t.src, x, t.si, t.sj);
            } else if (t.op == VX_OP_C_DIFF) {
                face(t.dst, x, t.di, t.dj);
                for (int k = 0; k < r; k++) face(t.src, x, x ? t.si : t.si + k, x ? t.sj + k : t.sj);
            }
        }
        if (!fix.empty()) {
            C.cut();
            for (const auto &c : fix)
                C.push(mk(VX_OP_CORRECT, std::get<0>(c), std::get<1>(c), std::get<2>(c), 0, 0, 0, 0));
        }
        pushcc_prog = C.upload(*dev);
    }
    for (auto &lp : levels) {
        Level &L = *lp;
        L.dev = dev.get();
        L.table = table;
        // coarse faces on a finer patch's edge (Rectangle.cpp:1313-1394)
        TaskList F;
        for (auto &K : L.rectangles) {
            for (int i = 0; i < K->n_x + 1; i++)
                for (int j = -1; j < K->n_p + 1; j++)
                    if (K->is_ilb_x[K->NS(i, j)]) {
                        Rectangle &Fp = *K->finer_level_x[K->NS(i, j)];
                        int fi = (K->x_pos + i) * r - Fp.x_pos, fj = (K->p_pos + j) * r - Fp.p_pos, aux = 0;
                        for (int k = 0; k < r; k++) aux |= (Fp.is_ilb_x[Fp.NS(fi, fj + k)] ? 1 : 0) << k;
                        F.tasks.push_back(mk(VX_OP_FLUX_X, K->table_index, i, j, Fp.table_index, fi, fj, aux));
                    }
            for (int i = -1; i < K->n_x + 1; i++)
                for (int j = 0; j < K->n_p + 1; j++)
                    if (K->is_ilb_p[K->NS(i, j)]) {
                        Rectangle &Fp = *K->finer_level_p[K->NS(i, j)];
                        int fi = (K->x_pos + i) * r - Fp.x_pos, fj = (K->p_pos + j) * r - Fp.p_pos, aux = 0;
                        for (int k = 0; k < r; k++) aux |= (Fp.is_ilb_p[Fp.NS(fi + k, fj)] ? 1 : 0) << k;
                        F.tasks.push_back(mk(VX_OP_FLUX_P, K->table_index, i, j, Fp.table_index, fi, fj, aux));
                    }
        }
        L.flux_prog = F.upload(*dev);  // faces are distinct: one batch
    }
}

// Mesh.cpp:116-130 with Rectangle.cpp:892-941, 1100-1128: the reference's
// per-cell transfer loops as rectangles over a table holding the new patches,
// then the old ones, one launch per phase (vx_transfer_ranges):
//   for each level i: the new patches interpolated from the old next-coarser
//   level, then copied from the old patches of level i (the overlaps);
//   the coarsest level copied from its old patches;
//   from the coarsest new level down, the finer cells no earlier phase wrote,
//   interpolated from the new next-coarser level.
// Each phase writes a cell at most once and reads only cells of old patches or
// of earlier phases, so the phases reproduce the reference's sequential order.
void Mesh::InterMeshDataTransfer(std::vector<std::unique_ptr<Level>> &old) {
    const int r = settings.refinementRatio;
    std::vector<Rectangle *> all = order;
    for (auto &l : old)
        for (auto &o : l->rectangles) {
            o->UnaliasSlots();
            o->table_index = (int)all.size();
            all.push_back(o.get());
        }
    std::vector<std::vector<vx_range_task>> phases;
    auto range = [](int op, Rectangle &D, Rectangle &S, int di0, int dj0, int si0, int sj0, int ni, int nj) {
        vx_range_task t{};
        t.op = op;
        t.dst = D.table_index;
        t.src = S.table_index;
        t.di0 = di0;
        t.dj0 = dj0;
        t.si0 = si0;
        t.sj0 = sj0;
        t.ni = ni;
        t.nj = nj;
        return t;
    };
    // coarse patch C interpolates the r x r fine cells of R under each coarse cell
    // (i, j) of C's range within one coarse cell of R; only_missing: the blocks no
    // earlier phase wrote, as runs along p
    auto from_coarse = [&](Rectangle &C, Rectangle &R, bool only_missing, std::vector<vx_range_task> &ph) {
        const int xp2 = R.x_pos / r, pp2 = R.p_pos / r, nx2 = R.n_x / r, np2 = R.n_p / r;
        const int lx = std::max(C.x_pos, xp2 - 1), ux = std::min(C.x_pos + C.n_x, xp2 + nx2 + 1);
        const int lp = std::max(C.p_pos, pp2 - 1), up_ = std::min(C.p_pos + C.n_p, pp2 + np2 + 1);
        if (ux <= lx || up_ <= lp) return;
        if (!only_missing) {
            ph.push_back(range(VX_OP_COARSE, R, C, xp2, pp2, lx, lp, ux - lx, up_ - lp));
            for (int i = lx; i < ux; i++)
                for (int j = lp; j < up_; j++)
                    for (int k = 0; k < r; k++)
                        for (int l = 0; l < r; l++) R.is_interpolated[R.NS((i - xp2) * r + k, (j - pp2) * r + l)] = 1;
            return;
        }
        for (int i = lx; i < ux; i++) {
            int j = lp;
            while (j < up_) {
                if (R.is_interpolated[R.NS(r * (i - xp2) + 1, r * (j - pp2) + 1)]) {
                    j++;
                    continue;
                }
                int j1 = j + 1;
                while (j1 < up_ && !R.is_interpolated[R.NS(r * (i - xp2) + 1, r * (j1 - pp2) + 1)]) j1++;
                ph.push_back(range(VX_OP_COARSE, R, C, xp2, pp2, i, j, 1, j1 - j));
                j = j1;
            }
        }
    };
    auto from_same = [&](Rectangle &O, Rectangle &R, std::vector<vx_range_task> &ph) {  // Rectangle.cpp:920-941
        const int lx = std::max(O.x_pos, R.x_pos - 2), ux = std::min(O.x_pos + O.n_x, R.x_pos + R.n_x + 2);
        const int lp = std::max(O.p_pos, R.p_pos - 2), up_ = std::min(O.p_pos + O.n_p, R.p_pos + R.n_p + 2);
        if (ux <= lx || up_ <= lp) return;
        ph.push_back(range(VX_OP_COPY, R, O, lx - R.x_pos, lp - R.p_pos, lx - O.x_pos, lp - O.p_pos, ux - lx,
                           up_ - lp));
        for (int i = lx; i < ux; i++)
            for (int j = lp; j < up_; j++) R.is_interpolated[R.NS(i - R.x_pos, j - R.p_pos)] = 1;
    };
    for (unsigned i = 0; i + 1 < old.size(); i++) {
        phases.emplace_back();
        for (auto &o : old[i + 1]->rectangles)
            for (auto &n : levels[i]->rectangles) from_coarse(*o, *n, false, phases.back());
        phases.emplace_back();
        for (auto &o : old[i]->rectangles)
            for (auto &n : levels[i]->rectangles) from_same(*o, *n, phases.back());
    }
    if (old.size() > 1) {
        int i = (int)old.size() - 1;
        phases.emplace_back();
        for (auto &o : old[i]->rectangles)
            for (auto &n : levels[i]->rectangles) from_same(*o, *n, phases.back());
    }
    // the missing cells: each phase reads the level the previous phase completed
    std::vector<vx_patch> d;
    for (Rectangle *x : all) d.push_back(x->desc());
    vx_patch_table *tb = nullptr;
    dev->check(vx_patch_table_create(dev->ctx, d.data(), (int32_t)d.size(), &tb), "transfer table");
    int rc = VX_OK;
    for (auto &ph : phases)
        if (rc == VX_OK) rc = vx_transfer_ranges(dev->ctx, tb, ph.data(), (int32_t)ph.size(), r);
    for (int i = (int)levels.size() - 1; i > 0 && rc == VX_OK; i--) {
        std::vector<vx_range_task> ph;
        for (auto &c : levels[i]->rectangles)
            for (auto &f : levels[i - 1]->rectangles) from_coarse(*c, *f, true, ph);
        rc = vx_transfer_ranges(dev->ctx, tb, ph.data(), (int32_t)ph.size(), r);
    }
    vx_patch_table_destroy(tb);
    dev->check(rc, "regrid transfer");
}

// SolverManager.cpp:48-58 (without the screen output)
void SolverManager::reGrid(double t) {
    DropGraphs();  // recorded steps hold the old patch tables and programs
    // SolverManager.cpp:49-51: the time in periods of the laser (tempEM[0] / c)
    double T = (!settings.tempEM.empty() && settings.tempEM[0] > 0) ? settings.tempEM[0] / vx_cs : 1.0;
    {
        std::ostringstream o;
        o << std::setfill(' ') << std::setw(4) << ' ' << std::fixed << std::setprecision(5) << t / T
          << std::setw(5) << ' ';
        ScreenWrite(o.str());
    }
    for (auto &m : meshes) {
        m->updateHierarchy();
        m->SetFieldSolver(EMSolver.get());
        ScreenRow(*m);
    }
    EMSolver->RebuildMomentPlan();
}

}  // namespace vxh
