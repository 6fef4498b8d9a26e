// Photon kd tree on the host (photon_map.h:83-114).  The tree depends only on
// the comparisons std::nth_element makes, so sorting (position, index) records
// yields the reference's permutation of its photon records; subtrees below the
// top levels are built on threads and spliced in pre-order.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "host.hpp"

namespace lrt {
namespace {

struct Rec {
    float p[3];
    uint32_t idx;
};

struct Node {
    uint32_t w[4];
};

constexpr int kLeafMax = 8;
constexpr uint32_t kParallelMin = 1u << 15;

struct Builder {
    std::vector<Rec>& r;
    int par_levels;

    int build(uint32_t start, uint32_t count, int level, std::vector<Node>& out, int& depth) {
        const int index = (int)out.size();
        out.push_back(Node{});
        depth = std::max(depth, level + 1);
        if (count <= (uint32_t)kLeafMax) {
            out[(size_t)index] = Node{{0u, 3u, start, count}};
            return index;
        }
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = r[start].p[a];
        for (uint32_t i = start + 1; i < start + count; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], r[i].p[a]);
                hi[a] = std::max(hi[a], r[i].p[a]);
            }
        const float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        int axis = 0;
        if (e[1] > e[axis]) axis = 1;
        if (e[2] > e[axis]) axis = 2;
        const uint32_t mid = start + count / 2;
        std::nth_element(r.begin() + start, r.begin() + mid, r.begin() + start + count,
                         [axis](const Rec& a, const Rec& b) { return a.p[axis] < b.p[axis]; });
        uint32_t left, right;
        if (level < par_levels && count >= kParallelMin) {
            std::vector<Node> L, R;
            int dl = 0, dr = 0;
            std::thread t([&] { build(start, mid - start, level + 1, L, dl); });
            build(mid, start + count - mid, level + 1, R, dr);
            t.join();
            depth = std::max(depth, std::max(dl, dr));
            left = (uint32_t)out.size();
            for (Node nd : L) {
                if (nd.w[1] != 3u) { nd.w[2] += left; nd.w[3] += left; }
                out.push_back(nd);
            }
            right = (uint32_t)out.size();
            for (Node nd : R) {
                if (nd.w[1] != 3u) { nd.w[2] += right; nd.w[3] += right; }
                out.push_back(nd);
            }
        } else {
            left = (uint32_t)build(start, mid - start, level + 1, out, depth);
            right = (uint32_t)build(mid, start + count - mid, level + 1, out, depth);
        }
        uint32_t split;
        std::memcpy(&split, &r[mid].p[axis], 4);  // after both subtrees, like the reference
        out[(size_t)index] = Node{{split, (uint32_t)axis, left, right}};
        return index;
    }
};

}  // namespace

void photon_kd_build(const float* xyz, uint32_t n, PhotonKd& out, int threads) {
    out.nodes.clear();
    out.order.clear();
    out.depth = 0;
    if (n == 0) return;
    std::vector<Rec> r(n);
    for (uint32_t i = 0; i < n; ++i) r[i] = Rec{{xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]}, i};
    const int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    int levels = 0;
    while ((1 << levels) < nt && levels < 4) ++levels;
    Builder b{r, levels};
    std::vector<Node> nodes;
    nodes.reserve(2 * (size_t)n / kLeafMax + 16);
    int depth = 0;
    b.build(0, n, 0, nodes, depth);
    out.nodes.resize(nodes.size() * 4);
    std::memcpy(out.nodes.data(), nodes.data(), nodes.size() * sizeof(Node));
    out.order.resize(n);
    for (uint32_t i = 0; i < n; ++i) out.order[i] = r[i].idx;
    out.depth = depth;
}

}  // namespace lrt
