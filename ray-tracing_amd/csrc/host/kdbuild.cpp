// SAH kd-tree build with the reference's split rules (sah_kdtree.h:35-57,
// 141-347) and output: same node numbering, same reference order, same
// retracted orphans.  What differs is the work schedule only:
//   * events are sorted ONCE at the root as packed 64-bit keys
//     (axis | order-preserving position bits | type | index) and every child
//     receives the stable sub-sequence of its primitives, so no per-node sort;
//     the reference sorts (axis, position, type, index), a total order, and a
//     sub-sequence of a sorted list is sorted, so the sweep sees the same order;
//   * the N-byte side array is per worker thread instead of per node (:290);
//   * large subtrees are built concurrently and spliced back in pre-order.
// -0.0 positions are folded into +0.0 (the reference compares positions with
// ==, which already treats them as equal).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <stdexcept>
#include <thread>

#include "host.hpp"

namespace lrt {
namespace {

constexpr float kTraversal = 15.0f, kQuery = 20.0f, kEmptyBonus = 0.9f;
constexpr int kStopPrims = 6, kMaxBad = 3;
constexpr uint32_t kParallelPrims = 16384;

inline uint32_t f2key(float f) {
    f = f + 0.0f;  // -0 -> +0
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float key2f(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint64_t make_ev(uint32_t axis, float pos, uint32_t type, uint32_t idx) {
    return ((uint64_t)axis << 62) | ((uint64_t)f2key(pos) << 30) | ((uint64_t)type << 28) | (uint64_t)idx;
}
inline uint32_t ev_axis(uint64_t e) { return (uint32_t)(e >> 62); }
inline uint32_t ev_posk(uint64_t e) { return (uint32_t)(e >> 30); }
inline uint32_t ev_type(uint64_t e) { return (uint32_t)(e >> 28) & 3u; }
inline uint32_t ev_idx(uint64_t e) { return (uint32_t)e & 0x0fffffffu; }

struct Box {
    float lo[3], hi[3];
    float area() const {
        const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        return 2.0f * (ex * ey + ex * ez + ey * ez);
    }
};

inline void split_probability(const Box& b, int axis, float lw, float rw, float& pl, float& pr) {
    const float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
    const float denom = ex * ey + ey * ez + ex * ez;
    const float inv = 1.0f / denom;
    float face, perim;
    if (axis == 0) { face = ey * ez; perim = ey + ez; }
    else if (axis == 1) { face = ex * ez; perim = ex + ez; }
    else { face = ex * ey; perim = ex + ey; }
    pl = (face + perim * lw) * inv;
    pr = (face + perim * rw) * inv;
}

struct Sub {
    std::vector<KdNode> nodes;
    std::vector<uint32_t> refs;
};

class Builder {
public:
    Builder(uint32_t n, int depth_limit, int threads) : n_(n), depth_limit_(depth_limit), budget_(threads - 1) {}

    float build(Sub& out, std::vector<uint32_t>& prims, std::vector<uint64_t>& ev, const Box& voxel, int depth,
                int bad, std::vector<uint8_t>& side) {
        const int index = (int)out.nodes.size();
        out.nodes.emplace_back();
        const int n = (int)prims.size();
        const float leaf_cost = (float)n * kQuery;
        if (n <= kStopPrims || depth >= depth_limit_ || voxel.area() <= 0.0f) {
            make_leaf(out, index, prims);
            return leaf_cost;
        }

        // event sweep (sah_kdtree.h:213-278)
        int nl[3] = {0, 0, 0}, nr[3] = {n, n, n};
        float best = std::numeric_limits<float>::infinity(), best_pos = 0.0f;
        int best_axis = 0;
        bool planar_left = false;
        size_t axis_begin[4] = {ev.size(), ev.size(), ev.size(), ev.size()};
        const size_t m = ev.size();
        for (size_t c = 0; c < m;) {
            const uint32_t axis = ev_axis(ev[c]);
            if (axis_begin[axis] == m) axis_begin[axis] = c;
            const uint32_t pk = ev_posk(ev[c]);
            const float pos = key2f(pk);
            int n_end = 0, n_planar = 0, n_start = 0;
            // group = same axis and position: the 34 high bits of the key
            const uint64_t group = ev[c] >> 30;
            while (c < m && (ev[c] >> 30) == group) {
                const uint32_t t = ev_type(ev[c]);
                n_end += t == 0;
                n_planar += t == 1;
                n_start += t == 2;
                ++c;
            }
            nr[axis] -= n_planar + n_end;
            if (pos > voxel.lo[axis] && pos < voxel.hi[axis]) {
                float pl, pr;
                split_probability(voxel, (int)axis, pos - voxel.lo[axis], voxel.hi[axis] - pos, pl, pr);
                const int a = nl[axis], b = nr[axis];
                if (n_planar == 0) {
                    float cost = kTraversal + kQuery * (pl * (float)a + pr * (float)b);
                    if (a == 0 || b == 0) cost *= kEmptyBonus;
                    if (cost < best) { best = cost; best_axis = (int)axis; best_pos = pos; }
                } else {
                    float cl = kTraversal + kQuery * (pl * (float)(a + n_planar) + pr * (float)b);
                    float cr = kTraversal + kQuery * (pl * (float)a + pr * (float)(b + n_planar));
                    if (a + n_planar == 0 || b == 0) cl *= kEmptyBonus;
                    if (a == 0 || b + n_planar == 0) cr *= kEmptyBonus;
                    if (cl < best || cr < best) {
                        planar_left = cl < cr;
                        best = planar_left ? cl : cr;
                        best_axis = (int)axis;
                        best_pos = pos;
                    }
                }
            }
            nl[axis] += n_start + n_planar;
        }
        for (int a = 2; a >= 0; --a)
            if (axis_begin[a] == m) axis_begin[a] = axis_begin[a + 1];

        if (best >= leaf_cost) {
            if ((best > 4.0f * leaf_cost && n < 16) || bad >= kMaxBad || best == std::numeric_limits<float>::infinity()) {
                make_leaf(out, index, prims);
                return leaf_cost;
            }
            ++bad;
        }

        // classification (sah_kdtree.h:290-309): 1 left, 2 right, 0 both
        for (uint32_t p : prims) side[p] = 0;
        for (size_t c = axis_begin[best_axis]; c < axis_begin[best_axis + 1]; ++c) {
            const uint64_t e = ev[c];
            const uint32_t t = ev_type(e);
            const float pos = key2f(ev_posk(e));
            if (t == 0 && pos <= best_pos) side[ev_idx(e)] = 1;
            else if (t == 2 && pos >= best_pos) side[ev_idx(e)] = 2;
            else if (t == 1) side[ev_idx(e)] = (pos < best_pos || (pos == best_pos && planar_left)) ? 1 : 2;
        }
        std::vector<uint32_t> lp, rp;
        lp.reserve(prims.size());
        rp.reserve(prims.size());
        for (uint32_t p : prims) {
            const uint8_t s = side[p];
            if (s != 2) lp.push_back(p);
            if (s != 1) rp.push_back(p);
        }
        std::vector<uint64_t> le, re;
        le.reserve(lp.size() * 6);
        re.reserve(rp.size() * 6);
        for (uint64_t e : ev) {
            const uint8_t s = side[ev_idx(e)];
            if (s != 2) le.push_back(e);
            if (s != 1) re.push_back(e);
        }
        std::vector<uint64_t>().swap(ev);

        Box lv = voxel, rv = voxel;
        lv.hi[best_axis] = best_pos;
        rv.lo[best_axis] = best_pos;

        float lc, rc;
        int li, ri;
        const bool split_work = n >= (int)kParallelPrims && take_thread();
        if (split_work) {
            Sub ls, rs;
            auto fut = std::async(std::launch::async, [&]() {
                std::vector<uint8_t> side2(n_, 0);
                const float c = build(ls, lp, le, lv, depth + 1, bad, side2);
                give_thread();
                return c;
            });
            rc = build(rs, rp, re, rv, depth + 1, bad, side);
            lc = fut.get();
            li = (int)out.nodes.size();
            splice(out, ls);
            ri = (int)out.nodes.size();
            splice(out, rs);
        } else {
            li = (int)out.nodes.size();
            lc = build(out, lp, le, lv, depth + 1, bad, side);
            ri = (int)out.nodes.size();
            rc = build(out, rp, re, rv, depth + 1, bad, side);
        }
        float pl, pr;
        split_probability(voxel, best_axis, best_pos - voxel.lo[best_axis], voxel.hi[best_axis] - best_pos, pl, pr);
        const float final_cost = kTraversal + pl * lc + pr * rc;
        if (final_cost < leaf_cost) {
            KdNode& nd = out.nodes[(size_t)index];
            nd.leaf = 0;
            nd.axis = (uint8_t)best_axis;
            nd.split = best_pos;
            nd.left = li;
            nd.right = ri;
            return final_cost;
        }
        make_leaf(out, index, prims);
        return leaf_cost;
    }

private:
    static void make_leaf(Sub& out, int index, const std::vector<uint32_t>& prims) {
        KdNode& nd = out.nodes[(size_t)index];
        nd.leaf = 1;
        nd.first = (int)out.refs.size();
        nd.count = (int)prims.size();
        out.refs.insert(out.refs.end(), prims.begin(), prims.end());
    }
    static void splice(Sub& out, const Sub& s) {
        const int no = (int)out.nodes.size(), ro = (int)out.refs.size();
        for (KdNode nd : s.nodes) {
            if (nd.leaf) nd.first += ro;
            else { nd.left += no; nd.right += no; }
            out.nodes.push_back(nd);
        }
        out.refs.insert(out.refs.end(), s.refs.begin(), s.refs.end());
    }
    bool take_thread() {
        int b = budget_.load();
        while (b > 0)
            if (budget_.compare_exchange_weak(b, b - 1)) return true;
        return false;
    }
    void give_thread() { budget_.fetch_add(1); }

    uint32_t n_;
    int depth_limit_;
    std::atomic<int> budget_;
};

}  // namespace

void kd_build(const float* boxes, uint32_t n, KdTree& out, int threads) {
    out.nodes.clear();
    out.refs.clear();
    if (n == 0) return;
    if (n > 0x0fffffffu) throw std::runtime_error("kd_build: more than 2^28 primitives");
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    Box root;
    for (int a = 0; a < 3; ++a) { root.lo[a] = INFINITY; root.hi[a] = -INFINITY; }
    std::vector<uint64_t> ev;
    ev.reserve((size_t)n * 6);
    for (uint32_t i = 0; i < n; ++i) {
        const float* b = boxes + 6 * (size_t)i;
        for (int a = 0; a < 3; ++a) {
            // aabb::extend(box) extends by min then max corner (aabb.h:17-20)
            root.lo[a] = std::min(root.lo[a], b[a]);
            root.hi[a] = std::max(root.hi[a], b[a]);
            root.lo[a] = std::min(root.lo[a], b[3 + a]);
            root.hi[a] = std::max(root.hi[a], b[3 + a]);
            if (b[a] == b[3 + a]) {
                ev.push_back(make_ev((uint32_t)a, b[a], 1, i));
            } else {
                ev.push_back(make_ev((uint32_t)a, b[a], 2, i));
                ev.push_back(make_ev((uint32_t)a, b[3 + a], 0, i));
            }
        }
    }
    std::sort(ev.begin(), ev.end());
    const int lim = std::min(48, (int)(8.0f + 1.3f * std::log2((float)n)));
    out.depth_limit = std::max(lim, 1);
    std::vector<uint32_t> prims(n);
    for (uint32_t i = 0; i < n; ++i) prims[i] = i;
    std::vector<uint8_t> side(n, 0);
    Builder b(n, out.depth_limit, threads);
    Sub s;
    b.build(s, prims, ev, root, 1, 0, side);
    out.nodes = std::move(s.nodes);
    out.refs = std::move(s.refs);
}

void kd_flatten(const KdTree& t, FlatTree& out) {
    out.words.clear();
    out.refs.clear();
    out.max_depth = 0;
    if (t.nodes.empty()) return;
    // pre-order walk from the root over live nodes; leaves' references are
    // copied in visit order so every leaf range is contiguous
    struct Item { int node, depth; };
    std::vector<Item> stack{{0, 1}};
    std::vector<int> new_index(t.nodes.size(), -1);
    std::vector<std::pair<int, int>> right_fix;  // (new parent index, old right child)
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        const KdNode& nd = t.nodes[(size_t)it.node];
        const int idx = (int)(out.words.size() / 2);
        new_index[(size_t)it.node] = idx;
        out.max_depth = std::max(out.max_depth, it.depth);
        if (nd.leaf) {
            const uint32_t first = (uint32_t)out.refs.size();
            out.refs.insert(out.refs.end(), t.refs.begin() + nd.first, t.refs.begin() + nd.first + nd.count);
            out.words.push_back(first);
            out.words.push_back(3u | ((uint32_t)nd.count << 2));
        } else {
            uint32_t sb;
            std::memcpy(&sb, &nd.split, 4);
            out.words.push_back(sb);
            out.words.push_back((uint32_t)nd.axis);  // right child patched below
            right_fix.emplace_back(idx, nd.right);
            stack.push_back({nd.right, it.depth + 1});
            stack.push_back({nd.left, it.depth + 1});
        }
    }
    for (const auto& f : right_fix) {
        const int r = new_index[(size_t)f.second];
        out.words[2 * (size_t)f.first + 1] |= (uint32_t)r << 2;
    }
}

}  // namespace lrt
