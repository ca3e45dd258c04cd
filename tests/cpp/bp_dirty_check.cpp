// The broadphase tree's changed-node tracking (BroadPhase::dirty), as the
// device pair query relies on it: one tree moves every proxy with
// move_proxy (the reference's MoveProxy); another is moved the way the
// device path moves it (tight boxes written into a separate image, only the
// leaves that left their fat boxes re-inserted on the host, the host's other
// tight boxes left stale), and the separate image is brought up to date from
// the changed nodes only.  Both images must be equal, word for word, after
// every step.
#include "host_geometry.h"
#include <cstdio>
#include <cstring>
#include <random>
using namespace kz;
static bool contains(const Aabb& a, const Aabb& o) {
    return a.mn.x <= o.mn.x && a.mx.x >= o.mx.x && a.mn.y <= o.mn.y && a.mx.y >= o.mx.y && a.mn.z <= o.mn.z && a.mx.z >= o.mx.z;
}
int main() {
    const uint32_t n = 3000;
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(0.f, 20.f);
    std::normal_distribution<float> J(0.f, 0.06f);
    std::vector<f3> pos(n);
    BroadPhase R, H;
    std::vector<uint32_t> pr(n), ph(n);
    auto box = [&](uint32_t i) { return Aabb{sub(pos[i], f3{0.3f,0.3f,0.3f}), add(pos[i], f3{0.3f,0.3f,0.3f})}; };
    for (uint32_t i = 0; i < n; ++i) { pos[i] = f3{U(rng), U(rng) * 0.3f, U(rng)}; pr[i] = R.create_proxy(box(i), i, 1, ~0u); ph[i] = H.create_proxy(box(i), i, 1, ~0u); }
    std::vector<float> S(16 * size_t(R.num_nodes()));
    R.image(S.data());
    H.clear_dirty();
    std::vector<float> fat_dev(6 * n);  // device copy of leaf fat boxes lives in S
    long total_dirty = 0, total_moved = 0;
    for (int step = 0; step < 400; ++step) {
        for (uint32_t i = 0; i < n; ++i) { pos[i].x += J(rng); pos[i].z += J(rng); }
        // reference semantics
        for (uint32_t i = 0; i < n; ++i) R.move_proxy(pr[i], box(i), f3{0.01f, 0, 0});
        // device: tight boxes into S, moved list from S's fat boxes
        std::vector<uint32_t> moved;
        for (uint32_t i = 0; i < n; ++i) {
            float* r = &S[16 * size_t(ph[i])];
            const Aabb t = box(i);
            r[8] = t.mn.x; r[9] = t.mn.y; r[10] = t.mn.z; r[12] = t.mx.x; r[13] = t.mx.y; r[14] = t.mx.z;
            const Aabb fat{f3{r[0], r[1], r[2]}, f3{r[4], r[5], r[6]}};
            if (!contains(fat, t)) moved.push_back(i);
        }
        H.clear_dirty();
        for (uint32_t i : moved) { H.set_tight(ph[i], box(i)); H.reinsert(ph[i], f3{0.01f, 0, 0}); }
        total_moved += moved.size();
        if (H.num_nodes() * 16 != S.size()) { printf("tree grew at step %d\n", step); return 1; }
        for (uint32_t id : H.dirty()) H.image(&S[16 * size_t(id)], id, id + 1);
        total_dirty += H.dirty().size();
        std::vector<float> ref(16 * size_t(R.num_nodes()));
        R.image(ref.data());
        if (R.root() != H.root()) { printf("root differs at %d\n", step); return 1; }
        if (std::memcmp(ref.data(), S.data(), ref.size() * 4) != 0) {
            for (size_t k = 0; k < ref.size(); ++k) if (std::memcmp(&ref[k], &S[k], 4)) { printf("step %d: node %zu word %zu differs\n", step, k / 16, k % 16); return 1; }
        }
    }
    printf("400 steps identical; moved %ld, dirty records %ld\n", total_moved, total_dirty);
}
