// Binned SAH BVH for the device.
//
// Same cost model as the reference builder (src/render/bvh.cpp:101-196: C = 1 + (A_L N_L + A_R N_R)
// / A_P against a leaf cost of N, split by binned centroids) with 16 bins instead of 12, subtrees
// built in parallel, and two device-driven limits: leaves never exceed 16 triangles and the depth
// never exceeds 30, so the traversal stack is bounded by 32 entries. Closest-hit results do not
// depend on the tree shape (nearest t wins, equal t goes to the lower scene-order index,
// bvh.cpp:273-276), so the device tree is free to differ from the reference's.
#include "mrg_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <future>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace mrg {
namespace {

constexpr int kBins = 16;

struct Box {
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool   empty() const { return lo[0] > hi[0]; }
    double area() const {
        if (empty()) return 0;
        const double x = double(hi[0]) - lo[0], y = double(hi[1]) - lo[1], z = double(hi[2]) - lo[2];
        return 2 * (x * y + y * z + z * x);
    }
};

struct Prim {
    Box   box;
    float c[3];
};

struct TNode {
    Box                    box;
    std::unique_ptr<TNode> l, r;
    uint32_t               first = 0, count = 0;  // leaf when l == nullptr
};

uint32_t ceil_log2(uint64_t v) {
    uint32_t k = 0;
    while ((uint64_t(1) << k) < v) ++k;
    return k;
}

struct Builder {
    const std::vector<Prim>& prims;
    std::vector<uint32_t>&   order;
    std::atomic<int>         spare_threads;
    std::atomic<uint64_t>    leaves{0};
    std::atomic<uint64_t>    max_depth{0};

    Builder(const std::vector<Prim>& p, std::vector<uint32_t>& o, int threads) : prims(p), order(o), spare_threads(threads - 1) {}

    void note_depth(uint64_t d) {
        uint64_t cur = max_depth.load();
        while (d > cur && !max_depth.compare_exchange_weak(cur, d)) {
        }
    }

    std::unique_ptr<TNode> make_leaf(const Box& b, uint32_t begin, uint32_t end, uint64_t depth) {
        auto n   = std::make_unique<TNode>();
        n->box   = b;
        n->first = begin;
        n->count = end - begin;
        ++leaves;
        note_depth(depth);
        return n;
    }

    std::unique_ptr<TNode> build(uint32_t begin, uint32_t end, uint64_t depth) {
        Box bounds, cb;
        for (uint32_t i = begin; i < end; ++i) {
            bounds.grow(prims[order[i]].box);
            cb.grow(prims[order[i]].c);
        }
        const uint32_t count = end - begin;
        if (count == 1 || (count <= 4 && depth > 0)) return make_leaf(bounds, begin, end, depth);
        const bool must_balance = depth + ceil_log2((count + 3) / 4) >= kMaxDepth;
        int        best_axis = -1;
        int        best_split = 0;
        double     best_cost  = count <= kMaxLeaf && depth > 0 ? double(count) : std::numeric_limits<double>::infinity();
        if (!must_balance) {
            const double parent = bounds.area();
            for (int axis = 0; axis < 3; ++axis) {
                const float lo = cb.lo[axis], hi = cb.hi[axis];
                if (!(hi > lo)) continue;
                const float scale = kBins / (hi - lo);
                Box         bb[kBins];
                uint32_t    bn[kBins] = {};
                for (uint32_t i = begin; i < end; ++i) {
                    const Prim& p = prims[order[i]];
                    const int   k = std::min(kBins - 1, int((p.c[axis] - lo) * scale));
                    ++bn[k];
                    bb[k].grow(p.box);
                }
                double   right_area[kBins];
                uint32_t right_n[kBins];
                Box      acc;
                uint32_t n = 0;
                for (int k = kBins - 1; k >= 0; --k) {
                    acc.grow(bb[k]);
                    n += bn[k];
                    right_area[k] = acc.area();
                    right_n[k]    = n;
                }
                Box      left;
                uint32_t ln = 0;
                for (int s = 0; s + 1 < kBins; ++s) {
                    left.grow(bb[s]);
                    ln += bn[s];
                    if (ln == 0 || right_n[s + 1] == 0) continue;
                    const double cost = 1 + (left.area() * ln + right_area[s + 1] * right_n[s + 1]) / parent;
                    if (cost < best_cost) {
                        best_cost  = cost;
                        best_axis  = axis;
                        best_split = s;
                    }
                }
            }
        }
        uint32_t mid;
        if (best_axis >= 0) {
            const float lo = cb.lo[best_axis], hi = cb.hi[best_axis];
            const float scale = kBins / (hi - lo);
            const auto  it = std::partition(order.begin() + begin, order.begin() + end, [&](uint32_t i) {
                return std::min(kBins - 1, int((prims[i].c[best_axis] - lo) * scale)) <= best_split;
            });
            mid = uint32_t(it - order.begin());
        } else if (count <= kMaxLeaf && depth > 0 && !must_balance) {
            return make_leaf(bounds, begin, end, depth);
        } else if (count <= 4 && depth > 0) {
            return make_leaf(bounds, begin, end, depth);
        } else {
            // Object median on the widest centroid axis (SAH found nothing, or the depth budget
            // requires a balanced subtree).
            int   axis = 0;
            float w    = -1;
            for (int a = 0; a < 3; ++a) {
                if (cb.hi[a] - cb.lo[a] > w) {
                    w    = cb.hi[a] - cb.lo[a];
                    axis = a;
                }
            }
            mid = begin + count / 2;
            std::nth_element(order.begin() + begin, order.begin() + mid, order.begin() + end, [&](uint32_t a, uint32_t b) {
                return prims[a].c[axis] < prims[b].c[axis] || (prims[a].c[axis] == prims[b].c[axis] && a < b);
            });
        }
        auto node = std::make_unique<TNode>();
        node->box = bounds;
        if (count > 65536 && spare_threads.fetch_sub(1) > 0) {
            auto left = std::async(std::launch::async, [&, begin, mid, depth] { return build(begin, mid, depth + 1); });
            node->r   = build(mid, end, depth + 1);
            node->l   = left.get();
            ++spare_threads;
        } else {
            if (count > 65536) ++spare_threads;
            node->l = build(begin, mid, depth + 1);
            node->r = build(mid, end, depth + 1);
        }
        return node;
    }
};

void set_child(GpuNode& g, int which, const Box& b, uint32_t ref) {
    if (which == 0) {
        g.lo0[0] = b.lo[0];
        g.lo0[1] = b.lo[1];
        g.lo0[2] = b.lo[2];
        g.hi0x   = b.hi[0];
        g.hi0yz[0] = b.hi[1];
        g.hi0yz[1] = b.hi[2];
        g.ref0   = ref;
    } else {
        g.lo1xy[0] = b.lo[0];
        g.lo1xy[1] = b.lo[1];
        g.lo1z     = b.lo[2];
        g.hi1[0]   = b.hi[0];
        g.hi1[1]   = b.hi[1];
        g.hi1[2]   = b.hi[2];
        g.ref1     = ref;
    }
}

// Depth-first flattening: an inner node's left subtree follows it directly.
uint32_t flatten(const TNode* n, std::vector<GpuNode>& out) {
    const uint32_t index = uint32_t(out.size());
    out.emplace_back();
    uint32_t refs[2];
    const TNode* kids[2] = {n->l.get(), n->r.get()};
    for (int k = 0; k < 2; ++k) {
        const TNode* c = kids[k];
        refs[k]        = c->l ? flatten(c, out) : leaf_ref(c->first, c->count);
    }
    GpuNode& g = out[index];
    g          = GpuNode{};
    set_child(g, 0, kids[0]->box, refs[0]);
    set_child(g, 1, kids[1]->box, refs[1]);
    return index;
}

}  // namespace

BvhBuildResult build_bvh(const float* positions, uint64_t count, unsigned threads) {
    if (count >= (uint64_t(1) << 27)) throw std::invalid_argument("scene exceeds 2^27 triangles");
    BvhBuildResult out;
    out.order.resize(count);
    std::iota(out.order.begin(), out.order.end(), 0u);
    if (count == 0) return out;
    std::vector<Prim> prims(count);
    for (uint64_t i = 0; i < count; ++i) {
        const float* p = positions + i * 9;
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(p[k * 3]) || !std::isfinite(p[k * 3 + 1]) || !std::isfinite(p[k * 3 + 2]))
                throw std::invalid_argument("Transformed triangle position is not finite");
            prims[i].box.grow(p + k * 3);
        }
        for (int a = 0; a < 3; ++a) prims[i].c[a] = (p[a] + p[3 + a] + p[6 + a]) * (1.0f / 3.0f);
    }
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    Builder b(prims, out.order, int(threads));
    std::unique_ptr<TNode> root = b.build(0, uint32_t(count), 0);
    out.nodes.reserve(count);
    if (!root->l) {
        // A single triangle: a root with one leaf child and one empty child.
        GpuNode g{};
        set_child(g, 0, root->box, leaf_ref(root->first, root->count));
        Box empty;
        set_child(g, 1, empty, kEmptyRef);
        out.nodes.push_back(g);
    } else {
        flatten(root.get(), out.nodes);
    }
    out.leaves    = b.leaves.load();
    out.max_depth = b.max_depth.load();
    return out;
}

}  // namespace mrg
