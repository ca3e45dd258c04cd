// host_geometry.h - host-side geometry of the product library: convex hull
// construction, mass properties, world bounds and the dynamic AABB tree
// broadphase.  These run on the CPU when bodies are created and once per step
// (broadphase), exactly as in the reference engine, and decide the vertex
// order of hulls and the order of broadphase pairs that the device path
// consumes - so they follow the reference's arithmetic and tie-breaking.
#pragma once

#include <stdint.h>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "kz_math.h"

namespace kz {

struct Aabb {
    f3 mn{kInf, kInf, kInf};
    f3 mx{-kInf, -kInf, -kInf};
};

// Built convex hull: vertices in the order ConvexHullShape::Create emits them
// and its triangles (src/Geometry/Shapes.cpp:331-471).
struct Hull {
    std::vector<f3> verts;
    std::vector<uint32_t> tris;  // 3 per triangle
    Aabb bounds;
    float bounding_radius = 0.0f;
};

// ConvexHullShape::Create; false when the reference would return null.
bool build_hull(const float* points, uint32_t n, Hull& out, uint32_t max_points = 256);

struct MassProps {
    float mass = 0.0f;
    m3 inertia{};
    f3 com{0.0f, 0.0f, 0.0f};
};

// Shape::GetMassProperties for sphere, box, capsule, hull; plane = none.
MassProps shape_mass(uint32_t type, const float* params, const Hull* hull, float density);
// Shape::GetBoundingRadius.
float shape_bounding_radius(uint32_t type, const float* params, const Hull* hull);
// Shape::GetLocalBounds.
Aabb shape_local_bounds(uint32_t type, const float* params, const Hull* hull);
// AABB::Transformed (src/Core/Math.cpp:176-189).
Aabb transform_aabb(const Aabb& a, f4q q, f3 t);
// BodyManager::RecomputeMassProperties (src/Body/BodyManager.cpp:63-99):
// inverse mass and inverse local inertia for a dynamic body.
void body_mass(uint32_t motion_type, uint32_t type, const float* params, const Hull* hull, float density,
               float mass_override, float& inv_mass, m3& inv_inertia_local);

// Dynamic AABB tree + layer-filtered pair list (src/BroadPhase/BroadPhase.cpp,
// include/Kizuri/BroadPhase/BroadPhase.h:61-104).
class BroadPhase {
public:
    uint32_t create_proxy(const Aabb& tight, uint32_t user, uint32_t layer, uint32_t mask);
    bool move_proxy(uint32_t id, const Aabb& tight, f3 displacement);
    // move_proxy in two halves: set_tight stores the leaf's tight box and
    // tells whether its fat box still holds it (nothing else to do); else
    // reinsert gives the leaf its new fat box and puts it back in the tree.
    // set_tight touches only that leaf, so it may run for many leaves at once.
    bool set_tight(uint32_t id, const Aabb& tight);
    void reinsert(uint32_t id, f3 displacement);
    void update_pairs();
    const std::vector<uint32_t>& pairs() const { return pairs_; }  // proxy A, proxy B, ...
    uint32_t user(uint32_t id) const { return nodes_[id].user; }
    // Whether `id` is a leaf of the tree carrying `user` (snapshot checks).
    bool is_proxy(uint32_t id, uint32_t user) const;
    // The tree as the device query reads it: per node fat box + child1, fat
    // box + child2, tight box + user, tight box + height (ints as float bits),
    // 16 words; the root; the proxy layers and masks by node id.
    void image(float* nodes16) const;
    void image(float* nodes16, uint32_t first, uint32_t end) const;  // nodes [first, end) at nodes16
    // The nodes whose image record changed since clear_dirty() (leaf tight
    // boxes stored by set_tight are not counted: they are the caller's to
    // track), for updating an image kept elsewhere.
    const std::vector<uint32_t>& dirty() const { return dirty_; }
    void clear_dirty();
    uint32_t num_nodes() const { return uint32_t(nodes_.size()); }
    uint32_t root() const { return root_; }
    const std::vector<uint32_t>& layers() const { return layers_; }
    const std::vector<uint32_t>& masks() const { return masks_; }
    // The whole tree as plain words (snapshots): node records, root, free
    // list, proxy layers and masks.  restore() returns false on a malformed
    // image and leaves the tree unchanged.
    void save(std::vector<uint32_t>& out) const;
    bool restore(const uint32_t* in, size_t n);

private:
    static constexpr uint32_t kNull = 0xFFFFFFFFu;
    struct Node {
        Aabb fat, tight;
        uint32_t parent = kNull, child1 = kNull, child2 = kNull, user = 0, next = kNull;
        int32_t height = -1;
        bool leaf() const { return child1 == kNull; }
    };
    uint32_t allocate();
    void free_node(uint32_t id);
    void touch(uint32_t id);
    std::vector<uint32_t> dirty_;
    std::vector<uint8_t> mark_;
    void insert_leaf(uint32_t leaf);
    void remove_leaf(uint32_t leaf);
    uint32_t balance(uint32_t a);

    std::vector<Node> nodes_;
    uint32_t root_ = kNull, free_ = kNull;
    std::vector<uint32_t> layers_, masks_, pairs_;
    struct Entry {
        uint32_t n1, n2;
    };
    std::vector<Entry> stack_;
    static constexpr size_t kParallelLeaves = 8192;  // smaller trees are walked on the calling thread
    static constexpr size_t kSplitEntries = 4096;    // entries handed to the threads
    void query_step(std::vector<Entry>& stack, std::vector<uint32_t>& out) const;
};

// A few persistent host threads for per-body loops: run(n, f) calls
// f(begin, end) over contiguous chunks of [0, n), one chunk on the calling
// thread, and returns when all are done.
class HostPool {
public:
    explicit HostPool(unsigned workers);
    ~HostPool();
    HostPool(const HostPool&) = delete;
    HostPool& operator=(const HostPool&) = delete;
    template <class F> void run(uint32_t n, F&& f) {
        const uint32_t parts = uint32_t(threads_.size()) + 1;
        if (parts == 1 || n < 4096) {
            f(0u, n);
            return;
        }
        std::function<void(uint32_t)> job = [&f, n, parts](uint32_t k) {
            f(uint32_t(uint64_t(n) * k / parts), uint32_t(uint64_t(n) * (k + 1) / parts));
        };
        dispatch(&job, parts - 1);
        job(0);
        wait();
    }

private:
    void dispatch(std::function<void(uint32_t)>* job, uint32_t workers);
    void wait();
    void loop(uint32_t k);
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::function<void(uint32_t)>* job_ = nullptr;
    uint32_t pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace kz
