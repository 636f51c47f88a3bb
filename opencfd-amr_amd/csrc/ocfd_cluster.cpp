// Fine-grid generation from tagged coarse cells (the grid_places step of an
// AMR regrid, which EBR's errorEst feeds: Source/Tagging.cpp:10-81):
// tagged cells are coarsened by the blocking factor, grouped by
// Berger-Rigoutsos signature clustering until every cluster reaches the grid
// efficiency (the cluster/cut rules of AMReX's ClusterList: hole cuts first,
// then the strongest inflection of the signature's second difference, then
// bisection; among equal kinds the direction whose smaller part is largest),
// intersected with the region the boxes may cover (the proper nesting
// domain of the level being refined), refined back by the blocking factor,
// joined where two boxes share a face
// (the BoxList::simplify step), and chopped to the maximum grid size
// (balanced pieces, AMReX BoxList::maxSize).  Host code; no device work.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "ocfd.h"
#include "ocfd_host_pool.hpp"

namespace ocfd {
void set_error(const std::string& msg);
}

namespace {

struct IBox {
    int lo[3], hi[3];
    long volume() const { return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1); }
};

struct Pt {
    int v[3];
};

enum CutKind { HOLE = 0, INFLECTION = 1, BISECT = 2, NONE = 3 };

// A cluster: a range [b, e) of a point array, the range's bounding box and
// its signatures (point counts per plane across each direction, over the
// box).  Chopping partitions the range in place, so clusters never copy
// points, and builds the two parts' boxes and signatures in the same pass.
struct Cluster {
    size_t b = 0, e = 0;
    IBox box{};
    std::vector<int> hist[3];
    // bounding box and signatures from a pass over the points
    void scan(const std::vector<Pt>& pts)
    {
        for (int d = 0; d < 3; ++d) {
            box.lo[d] = pts[b].v[d];
            box.hi[d] = pts[b].v[d];
        }
        for (size_t q = b; q < e; ++q)
            for (int d = 0; d < 3; ++d) {
                box.lo[d] = std::min(box.lo[d], pts[q].v[d]);
                box.hi[d] = std::max(box.hi[d], pts[q].v[d]);
            }
        for (int d = 0; d < 3; ++d) hist[d].assign(box.hi[d] - box.lo[d] + 1, 0);
        for (size_t q = b; q < e; ++q)
            for (int d = 0; d < 3; ++d) hist[d][pts[q].v[d] - box.lo[d]]++;
    }
    size_t size() const { return e - b; }
    double eff() const { return (double)size() / (double)box.volume(); }
};

// cut index along one direction from the signature hist[0..len-1] over [lo, lo+len-1]
int find_cut(const std::vector<int>& hist, int lo, CutKind& kind, std::vector<int>& d2)
{
    const int MINOFF = 2, CUT_THRESH = 2;
    kind = NONE;
    const int len = (int)hist.size();
    if (len <= 1) return lo;
    const int mid = len / 2;
    int cut = -1;
    for (int i = 0; i < len; ++i) {
        if (hist[i] == 0) {
            kind = HOLE;
            if (std::abs(cut - mid) > std::abs(i - mid)) {
                cut = i;
                if (i > mid) break;
            }
        }
    }
    if (kind == HOLE) return lo + cut;
    d2.assign(len, 0);
    for (int i = 1; i < len - 1; ++i) d2[i] = hist[i + 1] - 2 * hist[i] + hist[i - 1];
    int locmax = -1;
    for (int i = MINOFF; i < len - MINOFF; ++i) {
        const int prev = d2[i - 1], cur = d2[i];
        const int dif = std::abs(prev - cur);
        if (prev * cur < 0 && dif >= locmax) {
            if (dif > locmax) {
                kind = INFLECTION;
                cut = i;
                locmax = dif;
            } else if (std::abs(i - mid) < std::abs(cut - mid)) {
                cut = i;
            }
        }
    }
    if (locmax <= CUT_THRESH) {
        cut = mid;
        kind = BISECT;
    }
    return lo + cut;
}

// The two parts' boxes and signatures: counted over the parent's box during
// the partition pass, then trimmed to the planes that hold points (a box is
// the span of its nonzero signature entries in each direction).
void trim(Cluster& c, const IBox& parent, const int* const* h)
{
    for (int d = 0; d < 3; ++d) {
        const int n = parent.hi[d] - parent.lo[d] + 1;
        int a = 0, z = n - 1;
        while (a < z && h[d][a] == 0) ++a;
        while (z > a && h[d][z] == 0) --z;
        c.box.lo[d] = parent.lo[d] + a;
        c.box.hi[d] = parent.lo[d] + z;
        c.hist[d].assign(h[d] + a, h[d] + z + 1);
    }
}

// the cut plane of cluster c: direction and first index of the high part
void choose_cut(const Cluster& c, std::vector<int>& d2, int& dir, int& cd)
{
    CutKind kind[3];
    int cut[3];
    CutKind best = NONE;
    for (int d = 0; d < 3; ++d) {
        cut[d] = find_cut(c.hist[d], c.box.lo[d], kind[d], d2);
        best = std::min(best, kind[d]);
    }
    int minlen = -1;
    dir = -1;
    for (int d = 0; d < 3; ++d) {
        if (kind[d] != best) continue;
        const int m = std::min(cut[d] - c.box.lo[d], c.box.hi[d] - cut[d]);
        if (m >= minlen) {
            dir = d;
            minlen = m;
        }
    }
    cd = cut[dir];
}

// points of a cluster from which its chop partitions on the host threads
constexpr size_t kParChop = 16384;

// chop of a large cluster on the host threads: each thread counts its slice's
// low points and both parts' signatures, then scatters the slice into a
// scratch array at its offsets (low parts first, slice by slice), which is
// copied back.  The parts hold the same points as the serial chop's (in
// another order), so their boxes, signatures and every later cut are the same.
Cluster chop_parallel(Cluster& c, std::vector<Pt>& pts, int dir, int cd, int nt)
{
    const IBox parent = c.box;
    int n[3], off[3];
    size_t tot = 0;
    for (int d = 0; d < 3; ++d) {
        n[d] = parent.hi[d] - parent.lo[d] + 1;
        off[d] = (int)tot;
        tot += n[d];
    }
    const size_t N = c.e - c.b, per = (N + nt - 1) / nt;
    std::vector<std::vector<int>> h(nt);
    std::vector<size_t> nlo(nt, 0);
    ocfd::parallel_for(nt, [&](int t) {
        const size_t a = c.b + std::min(N, t * per), z = c.b + std::min(N, (t + 1) * per);
        h[t].assign(2 * tot, 0);
        int* lo_h = h[t].data();
        int* hi_h = h[t].data() + tot;
        size_t k = 0;
        for (size_t q = a; q < z; ++q) {
            const Pt& p = pts[q];
            int* hh = p.v[dir] < cd ? (++k, lo_h) : hi_h;
            for (int d = 0; d < 3; ++d) hh[off[d] + p.v[d] - parent.lo[d]]++;
        }
        nlo[t] = k;
    });
    size_t nl = 0;
    std::vector<size_t> lo_at(nt), hi_at(nt);
    for (int t = 0; t < nt; ++t) { lo_at[t] = nl; nl += nlo[t]; }
    size_t nh = nl;
    for (int t = 0; t < nt; ++t) {
        const size_t a = std::min(N, t * per), z = std::min(N, (t + 1) * per);
        hi_at[t] = nh;
        nh += (z - a) - nlo[t];
    }
    // scratch kept between calls, per calling thread (the workers use the
    // caller's through this reference)
    static thread_local std::vector<Pt> scratch;
    if (scratch.size() < N) scratch.resize(N);
    std::vector<Pt>& tmp = scratch;
    ocfd::parallel_for(nt, [&](int t) {
        const size_t a = c.b + std::min(N, t * per), z = c.b + std::min(N, (t + 1) * per);
        size_t l = lo_at[t], u = hi_at[t];
        for (size_t q = a; q < z; ++q) {
            if (pts[q].v[dir] < cd) tmp[l++] = pts[q];
            else tmp[u++] = pts[q];
        }
    });
    ocfd::parallel_for(nt, [&](int t) {
        const size_t a = std::min(N, t * per), z = std::min(N, (t + 1) * per);
        std::copy(tmp.begin() + a, tmp.begin() + z, pts.begin() + c.b + a);
    });
    std::vector<int> sum(2 * tot, 0);
    for (int t = 0; t < nt; ++t)
        for (size_t q = 0; q < 2 * tot; ++q) sum[q] += h[t][q];
    Cluster hi;
    hi.b = c.b + nl;
    hi.e = c.e;
    c.e = c.b + nl;
    const int* lh[3] = {sum.data() + off[0], sum.data() + off[1], sum.data() + off[2]};
    const int* hh[3] = {sum.data() + tot + off[0], sum.data() + tot + off[1], sum.data() + tot + off[2]};
    trim(c, parent, lh);
    trim(hi, parent, hh);
    return hi;
}

// split cluster c in two; returns the high part, c keeps the low part
Cluster chop(Cluster& c, std::vector<Pt>& pts, std::vector<int>& d2, std::vector<int>& work)
{
    int dir, cd;
    choose_cut(c, d2, dir, cd);
    const int nt = ocfd::host_threads();
    if (nt > 1 && c.e - c.b >= kParChop) return chop_parallel(c, pts, dir, cd, nt);
    const IBox parent = c.box;
    int n[3], off[3];
    size_t tot = 0;
    for (int d = 0; d < 3; ++d) {
        n[d] = parent.hi[d] - parent.lo[d] + 1;
        off[d] = (int)tot;
        tot += n[d];
    }
    work.assign(2 * tot, 0);
    int* lo_h = work.data();
    int* hi_h = work.data() + tot;
    // partition by the cut plane, counting each part's signatures on the way
    size_t i = c.b, j = c.e;
    while (true) {
        while (i < j && pts[i].v[dir] < cd) {
            for (int d = 0; d < 3; ++d) lo_h[off[d] + pts[i].v[d] - parent.lo[d]]++;
            ++i;
        }
        while (i < j && pts[j - 1].v[dir] >= cd) {
            --j;
            for (int d = 0; d < 3; ++d) hi_h[off[d] + pts[j].v[d] - parent.lo[d]]++;
        }
        if (i >= j) break;
        std::swap(pts[i], pts[j - 1]);
    }
    Cluster hi;
    hi.b = i;
    hi.e = c.e;
    c.e = i;
    const int* lh[3] = {lo_h + off[0], lo_h + off[1], lo_h + off[2]};
    const int* hh[3] = {hi_h + off[0], hi_h + off[1], hi_h + off[2]};
    trim(c, parent, lh);
    trim(hi, parent, hh);
    return hi;
}

// Berger-Rigoutsos: clusters (in list order) chopped until each reaches eff;
// a chopped cluster keeps its place and is checked again, its high part goes
// to the end of the list.  Every cluster is refined from its own points only
// (into its own range of pts), so once the list holds enough clusters still
// to be chopped, they are refined on the host threads, each into a list of
// its own, and the lists joined in order.  The set of final clusters is the
// same either way; the boxes built from them are sorted before use.
void refine_serial(std::vector<Cluster>& lst, std::vector<Pt>& pts, double eff, size_t stop_at,
                   std::vector<int>& d2, std::vector<int>& work, size_t& q)
{
    while (q < lst.size()) {
        if (lst.size() >= stop_at) return;
        if (lst[q].eff() < eff) {
            Cluster hi = chop(lst[q], pts, d2, work);
            lst.push_back(std::move(hi));
        } else {
            ++q;
        }
    }
}

void refine(std::vector<Cluster>& lst, std::vector<Pt>& pts, double eff)
{
    std::vector<int> d2, work;
    size_t q = 0;
    const int nt = ocfd::host_threads();
    // serial until there are two clusters per thread (small point sets stay serial); on 8 host
    // threads two per thread ran the 1024x256 double-Mach level-1 set 0.07 ms faster than four
    const size_t npts = lst.empty() ? 0 : lst[0].e - lst[0].b;
    const size_t stop_at = (nt > 1 && npts >= 4096) ? (size_t)2 * nt : (size_t)-1;
    refine_serial(lst, pts, eff, stop_at, d2, work, q);
    if (q >= lst.size()) return;
    std::vector<size_t> todo;
    for (size_t c = q; c < lst.size(); ++c)
        if (lst[c].eff() < eff) todo.push_back(c);
    std::vector<std::vector<Cluster>> sub(todo.size());
    ocfd::parallel_for((int)todo.size(), [&](int t) {
        std::vector<int> d2t, workt;
        size_t qt = 0;
        sub[t].push_back(std::move(lst[todo[t]]));
        refine_serial(sub[t], pts, eff, (size_t)-1, d2t, workt, qt);
    });
    std::vector<Cluster> out;
    out.reserve(lst.size());
    size_t k = 0;
    for (size_t c = 0; c < lst.size(); ++c) {
        if (k < todo.size() && todo[k] == c) {
            for (Cluster& x : sub[k]) out.push_back(std::move(x));
            ++k;
        } else {
            out.push_back(std::move(lst[c]));
        }
    }
    lst.swap(out);
}

// join boxes that share their extent in two directions and touch in the
// third, until no pair joins: each pass sorts the list by lower corner (z
// major, stable) and lets every box, in list order, absorb the boxes after it
// that it can join, in list order, continuing after the last one absorbed.
// A box after box a in the list has a lower corner no smaller than a's, so
// the only boxes a can absorb start at (a.hi0+1, a.lo1, a.lo2),
// (a.lo0, a.hi1+1, a.lo2) or (a.lo0, a.lo1, a.hi2+1): they are found through
// an index by lower corner instead of a scan of the rest of the list.
struct LoKey {
    int v[3];
    bool operator==(const LoKey& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};
struct LoHash {
    size_t operator()(const LoKey& k) const
    {
        return ((size_t)(unsigned)k.v[0] * 0x9E3779B97F4A7C15ull) ^ ((size_t)(unsigned)k.v[1] * 0xC2B2AE3D27D4EB4Full) ^
               ((size_t)(unsigned)k.v[2] * 0x165667B19E3779F9ull);
    }
};

void merge(std::vector<IBox>& bx)
{
    bool changed = true;
    while (changed) {
        changed = false;
        std::stable_sort(bx.begin(), bx.end(), [](const IBox& a, const IBox& b) {
            if (a.lo[2] != b.lo[2]) return a.lo[2] < b.lo[2];
            if (a.lo[1] != b.lo[1]) return a.lo[1] < b.lo[1];
            return a.lo[0] < b.lo[0];
        });
        const size_t n = bx.size();
        std::unordered_map<LoKey, std::vector<size_t>, LoHash> by_lo;
        by_lo.reserve(2 * n);
        for (size_t i = 0; i < n; ++i) by_lo[LoKey{{bx[i].lo[0], bx[i].lo[1], bx[i].lo[2]}}].push_back(i);
        std::vector<char> dead(n, 0);
        for (size_t i = 0; i < n; ++i) {
            if (dead[i]) continue;
            IBox& a = bx[i];
            size_t p = i;   // the scan has passed every index up to p
            for (;;) {
                size_t best = n;
                for (int d = 0; d < 3; ++d) {
                    LoKey key{{a.lo[0], a.lo[1], a.lo[2]}};
                    key.v[d] = a.hi[d] + 1;
                    auto it = by_lo.find(key);
                    if (it == by_lo.end()) continue;
                    const std::vector<size_t>& v = it->second;
                    for (auto q = std::upper_bound(v.begin(), v.end(), p); q != v.end() && *q < best; ++q) {
                        if (dead[*q]) continue;
                        const IBox& b = bx[*q];
                        bool same = true;
                        for (int e = 0; e < 3; ++e)
                            if (e != d) same = same && b.hi[e] == a.hi[e];
                        if (same) { best = *q; break; }
                    }
                }
                if (best == n) break;
                for (int d = 0; d < 3; ++d) a.hi[d] = std::max(a.hi[d], bx[best].hi[d]);
                dead[best] = 1;
                changed = true;
                p = best;
            }
        }
        size_t w = 0;
        for (size_t i = 0; i < n; ++i)
            if (!dead[i]) bx[w++] = bx[i];
        bx.resize(w);
    }
}

// chop every box of the list to at most `chunk` cells per direction, in
// balanced pieces taken from the high end
void max_size(std::vector<IBox>& boxes, int chunk)
{
    for (int d = 0; d < 3; ++d) {
        std::vector<IBox> out;
        for (IBox b : boxes) {
            const int len = b.hi[d] - b.lo[d] + 1;
            if (len > chunk) {
                int ratio = 1, bs = chunk, nlen = len;
                while (bs % 2 == 0 && nlen % 2 == 0) {
                    ratio *= 2;
                    bs /= 2;
                    nlen /= 2;
                }
                const int numblk = nlen / bs + (nlen % bs ? 1 : 0);
                const int sz = nlen / numblk, extra = nlen % numblk;
                for (int k = 0; k < numblk - 1; ++k) {
                    const int ksize = (k < extra ? sz + 1 : sz) * ratio;
                    IBox top = b;
                    top.lo[d] = b.hi[d] - ksize + 1;
                    b.hi[d] = top.lo[d] - 1;
                    out.push_back(top);
                }
            }
            out.push_back(b);
        }
        boxes.swap(out);
    }
}


// floor division (box corners may be negative)
int fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// The clustering shared by both entry points: points (blocking-factor
// coordinates) grouped by Berger-Rigoutsos to grid_eff; when has_mask, each
// cluster intersected with the allowed region (ClusterList::intersect with the
// proper nesting domain: a box partly outside is replaced by an exact box
// cover of its part inside; in_box(box, out) returns true when every point of
// the box is allowed, else lists the allowed points in out); refined back by
// the blocking factor, joined, chopped to max size.
template <class InBox>
void cluster_points(std::vector<Pt>& all, bool has_mask, InBox&& in_box, int block, double grid_eff,
                    int max_size_cells, std::vector<IBox>& boxes)
{
    boxes.clear();
    if (all.empty()) return;
    std::vector<Cluster> lst(1);
    lst[0].b = 0;
    lst[0].e = all.size();
    lst[0].scan(all);
    refine(lst, all, grid_eff);
    std::vector<IBox> kept;
    if (!has_mask) {
        for (const Cluster& c : lst) kept.push_back(c.box);
    } else {
        // each cluster cut back to the allowed region on its own (host
        // threads), the pieces joined in cluster order
        std::vector<std::vector<IBox>> pieces(lst.size());
        ocfd::parallel_for((int)lst.size(), [&](int q) {
            const Cluster& c = lst[q];
            std::vector<Pt> in;
            if (in_box(c.box, in)) { pieces[q].push_back(c.box); return; }
            if (in.empty()) return;
            std::vector<Cluster> part(1);
            part[0].b = 0;
            part[0].e = in.size();
            part[0].scan(in);
            refine(part, in, 1.0);
            for (const Cluster& p : part) pieces[q].push_back(p.box);
        });
        for (const auto& v : pieces) kept.insert(kept.end(), v.begin(), v.end());
    }
    for (const IBox& kb : kept) {
        IBox b = kb;
        for (int d = 0; d < 3; ++d) {
            b.lo[d] *= block;
            b.hi[d] = (b.hi[d] + 1) * block - 1;
        }
        boxes.push_back(b);
    }
    merge(boxes);
    max_size(boxes, max_size_cells);
}

int write_boxes(const std::vector<IBox>& boxes, int cap, int* boxes_out, int* nbox_out, const char* who)
{
    *nbox_out = (int)boxes.size();
    if ((int)boxes.size() > cap) {
        ocfd::set_error(std::string(who) + ": more boxes than the output capacity");
        return OCFD_EINVAL;
    }
    for (size_t b = 0; b < boxes.size(); ++b)
        for (int d = 0; d < 3; ++d) {
            boxes_out[6 * b + d] = boxes[b].lo[d];
            boxes_out[6 * b + 3 + d] = boxes[b].hi[d];
        }
    return OCFD_OK;
}

} // namespace

extern "C" int ocfd_cluster_tags(const int lo[3], const int hi[3], const unsigned char* tags, const unsigned char* mask,
                                 int block, double grid_eff, int max_size_cells, int cap, int* boxes_out, int* nbox_out)
{
    using ocfd::set_error;
    if (!lo || !hi || !tags || !nbox_out || block < 1 || max_size_cells < block || max_size_cells % block ||
        grid_eff <= 0.0 || grid_eff > 1.0 || cap < 0 || (cap > 0 && !boxes_out)) {
        set_error("ocfd_cluster_tags: invalid arguments");
        return OCFD_EINVAL;
    }
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = hi[d] - lo[d] + 1;
        if (n[d] <= 0 || (lo[d] % block) || (n[d] % block)) {
            set_error("ocfd_cluster_tags: the tag box must be non-empty and aligned to the blocking factor");
            return OCFD_EINVAL;
        }
    }
    // tagged points at blocking-factor resolution (a block is tagged when any
    // of its cells is); the allowed region likewise, where a block counts only
    // when all its cells are allowed
    const int nb[3] = {n[0] / block, n[1] / block, n[2] / block};
    const size_t nblk = (size_t)nb[0] * nb[1] * nb[2];
    std::vector<unsigned char> bt_store, bm_store;
    const unsigned char* bt = tags;
    const unsigned char* bmp = mask;
    if (block > 1) {
        bt_store.assign(nblk, 0);
        if (mask) bm_store.assign(nblk, 1);
        size_t c = 0;
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j) {
                const size_t row = ((size_t)(k / block) * nb[1] + j / block) * nb[0];
                for (int i = 0; i < n[0]; ++i, ++c) {
                    if (tags[c]) bt_store[row + i / block] = 1;
                    if (mask && !mask[c]) bm_store[row + i / block] = 0;
                }
            }
        bt = bt_store.data();
        bmp = mask ? bm_store.data() : nullptr;
    }
    std::vector<Pt> all;
    {
        // tagged points, skipping untagged runs eight bytes at a time
        const size_t nrow = (size_t)nb[1] * nb[2];
        for (size_t r = 0; r < nrow; ++r) {
            const unsigned char* row = bt + r * nb[0];
            const int j = (int)(r % nb[1]), k = (int)(r / nb[1]);
            int i = 0;
            while (i < nb[0]) {
                if (i + 8 <= nb[0]) {
                    uint64_t w;
                    std::memcpy(&w, row + i, 8);
                    if (!w) { i += 8; continue; }
                }
                if (row[i]) all.push_back(Pt{{lo[0] / block + i, lo[1] / block + j, lo[2] / block + k}});
                ++i;
            }
        }
    }
    auto in_box = [&](const IBox& b, std::vector<Pt>& out) {
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
            for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                    const int ii = i - lo[0] / block, jj = j - lo[1] / block, kk = k - lo[2] / block;
                    if (bmp[((size_t)kk * nb[1] + jj) * nb[0] + ii]) out.push_back(Pt{{i, j, k}});
                }
        return (long)out.size() == b.volume();
    };
    std::vector<IBox> boxes;
    cluster_points(all, mask != nullptr, in_box, block, grid_eff, max_size_cells, boxes);
    return write_boxes(boxes, cap, boxes_out, nbox_out, "ocfd_cluster_tags");
}

extern "C" int ocfd_cluster_box_tags(int nbox, const int* lo, const int* hi, const unsigned char* flags_in,
                                     int nextra, const int* elo, const int* ehi, int use_mask, int block,
                                     double grid_eff, int max_size_cells, int cap, int* boxes_out, int* nbox_out)
{
    using ocfd::set_error;
    if (nbox < 0 || (nbox > 0 && (!lo || !hi || !flags_in)) || nextra < 0 || (nextra > 0 && (!elo || !ehi)) ||
        !nbox_out || block < 1 || max_size_cells < block || max_size_cells % block || grid_eff <= 0.0 ||
        grid_eff > 1.0 || cap < 0 || (cap > 0 && !boxes_out)) {
        set_error("ocfd_cluster_box_tags: invalid arguments");
        return OCFD_EINVAL;
    }
    std::vector<IBox> lv(nbox);
    std::vector<size_t> off(nbox + 1, 0);
    for (int b = 0; b < nbox; ++b) {
        for (int d = 0; d < 3; ++d) {
            lv[b].lo[d] = lo[3 * b + d];
            lv[b].hi[d] = hi[3 * b + d];
            if (lv[b].hi[d] < lv[b].lo[d]) {
                set_error("ocfd_cluster_box_tags: empty box");
                return OCFD_EINVAL;
            }
        }
        off[b + 1] = off[b] + (size_t)lv[b].volume();
    }
    // the extra boxes' cells count as tagged (tags |= their coverage)
    std::vector<unsigned char> fcopy;
    const unsigned char* flags = flags_in;
    if (nextra > 0) {
        fcopy.assign(flags_in, flags_in + off[nbox]);
        for (int e = 0; e < nextra; ++e)
            for (int b = 0; b < nbox; ++b) {
                const IBox& B = lv[b];
                int a[3], z[3];
                bool hit = true;
                for (int d = 0; d < 3; ++d) {
                    a[d] = std::max(B.lo[d], elo[3 * e + d]);
                    z[d] = std::min(B.hi[d], ehi[3 * e + d]);
                    hit = hit && a[d] <= z[d];
                }
                if (!hit) continue;
                const int n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1;
                for (int k = a[2]; k <= z[2]; ++k)
                    for (int j = a[1]; j <= z[1]; ++j) {
                        unsigned char* f = fcopy.data() + off[b] + ((size_t)(k - B.lo[2]) * n1 + (j - B.lo[1])) * n0;
                        for (int i = a[0]; i <= z[0]; ++i) f[i - B.lo[0]] |= 1;
                    }
            }
        flags = fcopy.data();
    }
    // points: blocks holding a cell flagged both tagged and allowed, in box
    // order; the level boxes are counted, then listed at their offsets, on
    // the host threads
    std::vector<size_t> at(nbox + 1, 0);
    ocfd::parallel_for(nbox, [&](int b) {
        const unsigned char* f = flags + off[b];
        size_t k = 0;
        for (size_t c = 0; c < off[b + 1] - off[b]; ++c) k += (f[c] & 3) == 3;
        at[b + 1] = k;
    });
    for (int b = 0; b < nbox; ++b) at[b + 1] += at[b];
    std::vector<Pt> all(at[nbox]);
    ocfd::parallel_for(nbox, [&](int b) {
        const IBox& B = lv[b];
        const unsigned char* f = flags + off[b];
        const int n0 = B.hi[0] - B.lo[0] + 1;
        Pt* out = all.data() + at[b];
        size_t c = 0;
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j, c += n0) {
                if (block == 1) {
                    for (int i = 0; i < n0; ++i)
                        if ((f[c + i] & 3) == 3) *out++ = Pt{{B.lo[0] + i, j, k}};
                    continue;
                }
                for (int i = 0; i < n0; ++i)
                    if ((f[c + i] & 3) == 3) *out++ = Pt{{fdiv(B.lo[0] + i, block), fdiv(j, block), fdiv(k, block)}};
            }
    });
    if (block > 1) {
        std::sort(all.begin(), all.end(), [](const Pt& a, const Pt& b) {
            return std::lexicographical_compare(a.v, a.v + 3, b.v, b.v + 3);
        });
        all.erase(std::unique(all.begin(), all.end(),
                              [](const Pt& a, const Pt& b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2]; }),
                  all.end());
    }
    // allowed points of a cluster box: blocks whose cells all lie in the
    // level and carry the allowed flag (called from several host threads)
    auto in_box = [&](const IBox& cb, std::vector<Pt>& out) {
        std::vector<unsigned char> loc;
        IBox cr;
        for (int d = 0; d < 3; ++d) {
            cr.lo[d] = cb.lo[d] * block;
            cr.hi[d] = (cb.hi[d] + 1) * block - 1;
        }
        const int m0 = cr.hi[0] - cr.lo[0] + 1, m1 = cr.hi[1] - cr.lo[1] + 1;
        loc.assign((size_t)cr.volume(), 0);
        for (int b = 0; b < nbox; ++b) {
            const IBox& B = lv[b];
            int a[3], z[3];
            bool hit = true;
            for (int d = 0; d < 3; ++d) {
                a[d] = std::max(B.lo[d], cr.lo[d]);
                z[d] = std::min(B.hi[d], cr.hi[d]);
                hit = hit && a[d] <= z[d];
            }
            if (!hit) continue;
            const int n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1;
            const unsigned char* f = flags + off[b];
            for (int k = a[2]; k <= z[2]; ++k)
                for (int j = a[1]; j <= z[1]; ++j) {
                    const unsigned char* src = f + ((size_t)(k - B.lo[2]) * n1 + (j - B.lo[1])) * n0 + (a[0] - B.lo[0]);
                    unsigned char* dst = loc.data() + ((size_t)(k - cr.lo[2]) * m1 + (j - cr.lo[1])) * m0 + (a[0] - cr.lo[0]);
                    for (int i = 0; i <= z[0] - a[0]; ++i) dst[i] = (src[i] >> 1) & 1;
                }
        }
        if (block == 1) {
            // the common case: every cell allowed, counted without listing
            bool all_ok = true;
            for (size_t q = 0; q < loc.size() && all_ok; ++q) all_ok = loc[q] != 0;
            if (all_ok) return true;
        }
        for (int k = cb.lo[2]; k <= cb.hi[2]; ++k)
            for (int j = cb.lo[1]; j <= cb.hi[1]; ++j)
                for (int i = cb.lo[0]; i <= cb.hi[0]; ++i) {
                    bool ok = true;
                    for (int kk = 0; kk < block && ok; ++kk)
                        for (int jj = 0; jj < block && ok; ++jj)
                            for (int ii = 0; ii < block && ok; ++ii)
                                ok = loc[((size_t)(k * block + kk - cr.lo[2]) * m1 + (j * block + jj - cr.lo[1])) * m0 +
                                         (i * block + ii - cr.lo[0])] != 0;
                    if (ok) out.push_back(Pt{{i, j, k}});
                }
        return (long)out.size() == cb.volume();
    };
    std::vector<IBox> boxes;
    cluster_points(all, use_mask != 0, in_box, block, grid_eff, max_size_cells, boxes);
    return write_boxes(boxes, cap, boxes_out, nbox_out, "ocfd_cluster_box_tags");
}

// Boxes of level f+1 projected two levels down: coarsened to level f, grown by
// n_proper there (the parts leaving the domain of level f wrapped to their
// periodic images on periodic sides, clipped elsewhere), coarsened to level
// f-1.  The tags Amr::grid_places adds so that level f's new grids keep the
// finer new grids properly nested.  The output boxes may overlap.
extern "C" int ocfd_project_boxes(int nbox, const int* lo, const int* hi, const int dom_lo[3], const int dom_hi[3],
                                  const int periodic[3], int n_proper, int cap, int* boxes_out, int* nbox_out)
{
    if (nbox < 0 || (nbox > 0 && (!lo || !hi)) || !dom_lo || !dom_hi || !periodic || n_proper < 0 || cap < 0 ||
        (cap > 0 && !boxes_out) || !nbox_out) {
        ocfd::set_error("ocfd_project_boxes: invalid arguments");
        return OCFD_EINVAL;
    }
    std::vector<IBox> out, parts, next;
    for (int b = 0; b < nbox; ++b) {
        IBox g;
        for (int d = 0; d < 3; ++d) {
            g.lo[d] = fdiv(lo[3 * b + d], 2) - n_proper;
            g.hi[d] = fdiv(hi[3 * b + d], 2) + n_proper;
        }
        parts.assign(1, g);
        for (int d = 0; d < 3; ++d) {
            const int n = dom_hi[d] - dom_lo[d] + 1;
            next.clear();
            for (const IBox& q : parts) {
                const int shifts[3] = {-n, 0, n};
                const int ns = periodic[d] ? 3 : 1;
                for (int t = 0; t < ns; ++t) {
                    const int sh = periodic[d] ? shifts[t] : 0;
                    const int a = std::max(q.lo[d] + sh, dom_lo[d]), z = std::min(q.hi[d] + sh, dom_hi[d]);
                    if (a > z) continue;
                    IBox r = q;
                    r.lo[d] = a;
                    r.hi[d] = z;
                    next.push_back(r);
                }
            }
            parts.swap(next);
        }
        for (IBox r : parts) {
            for (int d = 0; d < 3; ++d) {
                r.lo[d] = fdiv(r.lo[d], 2);
                r.hi[d] = fdiv(r.hi[d], 2);
            }
            out.push_back(r);
        }
    }
    return write_boxes(out, cap, boxes_out, nbox_out, "ocfd_project_boxes");
}

// Dense mask grown by n cells in every direction (Chebyshev distance, the
// TagBoxArray::buffer of amr.n_error_buf), or shrunk by n cells (the proper
// nesting domain of a level: cells whose n-neighbourhood the level covers).
// Separable: one windowed OR per axis.  Periodic axes wrap; outside the
// domain elsewhere counts as unset when growing and as set when shrinking.
extern "C" int ocfd_mask_grow(unsigned char* m, int nx, int ny, int nz, int n, const int* periodic, int shrink)
{
    if (!m || nx <= 0 || ny <= 0 || nz <= 0 || n < 0 || !periodic) {
        ocfd::set_error("ocfd_mask_grow: invalid arguments");
        return OCFD_EINVAL;
    }
    if (n == 0) return OCFD_OK;
    const size_t plane = (size_t)nx * ny, vol = plane * nz;
    if (shrink)
        for (size_t c = 0; c < vol; ++c) m[c] = !m[c];
    else
        for (size_t c = 0; c < vol; ++c) m[c] = m[c] != 0;
    // window [i-n, i+n] of one axis of length len, lines `stride` apart
    // (element step `step` along the axis), for `nl` lines at bases base(l)
    std::vector<unsigned char> tmp;
    auto pass = [&](int len, bool per, size_t step, auto&& line_base, size_t nl, size_t line_len) {
        // line_len elements move together (a row or a plane): OR whole rows
        tmp.resize((size_t)len * line_len);
        for (size_t l = 0; l < nl; ++l) {
            unsigned char* b = m + line_base(l);
            for (int a = 0; a < len; ++a) std::copy(b + a * step, b + a * step + line_len, tmp.data() + (size_t)a * line_len);
            for (int a = 0; a < len; ++a) {
                unsigned char* o = b + a * step;
                for (int s = -n; s <= n; ++s) {
                    if (s == 0) continue;
                    int q = a + s;
                    if (q < 0 || q >= len) {
                        if (!per || len == 1) continue;
                        q = ((q % len) + len) % len;
                    }
                    const unsigned char* src = tmp.data() + (size_t)q * line_len;
                    for (size_t e = 0; e < line_len; ++e) o[e] |= src[e];
                }
            }
        }
    };
    // x: each row on its own, as shifted ORs of the row's copy
    {
        std::vector<unsigned char> row((size_t)nx);
        const bool per = periodic[0] != 0 && nx > 1;
        for (size_t l = 0; l < (size_t)ny * nz; ++l) {
            unsigned char* o = m + l * nx;
            std::copy(o, o + nx, row.data());
            for (int s = -n; s <= n; ++s) {
                if (s == 0) continue;
                // o[i] |= row[i + s] where i + s lies inside the row
                const int a0 = std::max(0, -s), a1 = std::min(nx, nx - s);
                for (int i = a0; i < a1; ++i) o[i] |= row[i + s];
                if (per)
                    for (int i = 0; i < nx; ++i)
                        if (i + s < 0 || i + s >= nx) o[i] |= row[(((i + s) % nx) + nx) % nx];
            }
        }
    }
    // y: rows of a plane move together
    pass(ny, periodic[1] != 0, nx, [&](size_t l) { return l * plane; }, (size_t)nz, (size_t)nx);
    // z: whole planes move together
    pass(nz, periodic[2] != 0, plane, [&](size_t) { return (size_t)0; }, 1, plane);
    if (shrink)
        for (size_t c = 0; c < vol; ++c) m[c] = !m[c];
    return OCFD_OK;
}
