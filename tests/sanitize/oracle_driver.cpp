// Runs every entry point of the CPU oracle on small seeded inputs, for the AddressSanitizer /
// UndefinedBehaviorSanitizer build made by tests/test_host_sanitize.py: clouds of 1, 2, 17 and
// 3000 points (with duplicates and a collinear run), leaf sizes 1, 3 and 64, k above n, pixel
// lists and whole images, every exported array. Prints "oracle: ok" when all calls returned.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "oracle.h"

int main() {
    std::mt19937 g(7);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    for (int64_t n : {1, 2, 17, 3000}) {
        std::vector<float> p((size_t)(4 * n));
        for (int64_t i = 0; i < n; ++i) {
            p[4 * i] = u(g);
            p[4 * i + 1] = u(g);
            p[4 * i + 2] = i % 5 == 0 ? 0.0f : u(g);   // a plane of points
            p[4 * i + 3] = 1.0f;
            if (i % 7 == 3 && i > 0)                    // duplicates
                for (int c = 0; c < 3; ++c) p[4 * i + c] = p[4 * (i - 1) + c];
        }
        for (int leaf : {1, 3, 64}) {
            void* t = orc_tree_build(p.data(), n, leaf);
            if (!t) return 2;
            const int64_t nn = orc_tree_num_nodes(t);
            std::vector<int32_t> c1((size_t)nn), par((size_t)nn), b((size_t)nn), e((size_t)nn),
                order((size_t)n);
            std::vector<float> lo((size_t)(3 * nn)), hi((size_t)(3 * nn));
            orc_tree_export(t, c1.data(), par.data(), b.data(), e.data(), lo.data(), hi.data(),
                            order.data());
            (void)orc_tree_depth(t);
            const int m = 50;
            std::vector<float> q((size_t)(4 * m));
            for (float& v : q) v = 1.5f * u(g);
            for (int k : {1, 2, 10, 40}) {
                std::vector<int32_t> idx((size_t)(m * k));
                std::vector<float> d2((size_t)(m * k));
                orc_counters cnt{};
                orc_knn(t, q.data(), m, k, idx.data(), d2.data(), &cnt, 2);
                orc_knn_brute(p.data(), n, q.data(), m, k, idx.data(), d2.data(), 2);
            }
            std::vector<float> ru((size_t)n), rc((size_t)n);
            orc_counters cnt{};
            orc_radii(t, 0.1f, ru.data(), rc.data(), &cnt, 2);
            std::vector<float> col((size_t)(4 * n), 0.5f);
            const float pose[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0.1f, 0.2f, 3.0f, 1};
            const int W = 9, H = 7;
            const float pane = orc_pane_distance(W, 1.2f);
            std::vector<float> rgba((size_t)(4 * W * H));
            std::vector<int32_t> hit((size_t)(W * H)), steps((size_t)(W * H));
            orc_render(t, rc.data(), col.data(), pose, pane, W, H, nullptr, 0, rgba.data(),
                       hit.data(), &cnt, 2);
            const int32_t pix[3] = {0, 31, W * H - 1};
            orc_render_steps(t, rc.data(), col.data(), pose, pane, W, H, pix, 3, rgba.data(),
                             hit.data(), steps.data(), &cnt, 2);
            std::vector<uint8_t> ins((size_t)n);
            orc_frustum_inside(pose, pane, W, H, p.data(), n, ins.data());
            std::vector<double> x((size_t)(3 * n)), r((size_t)n, 0.05), c((size_t)(3 * n), 0.5);
            for (int64_t i = 0; i < n; ++i)
                for (int a = 0; a < 3; ++a) x[(size_t)(3 * i + a)] = p[(size_t)(4 * i + a)];
            std::vector<float> trgb((size_t)(3 * W * H), 0.3f), tdep((size_t)(W * H), 2.0f);
            std::vector<double> gx((size_t)(3 * n)), gr((size_t)n), gc((size_t)(3 * n));
            orc_render(t, rc.data(), col.data(), pose, pane, W, H, nullptr, 0, rgba.data(),
                       hit.data(), &cnt, 1);
            (void)orc_loss_grad_fixed(x.data(), r.data(), c.data(), n, pose, pane, W, H,
                                      hit.data(), trgb.data(), tdep.data(), 1.0, 0.5, gx.data(),
                                      gr.data(), gc.data());
            orc_tree_free(t);
        }
        std::vector<int32_t> sub((size_t)n);
        orc_subsample_indices(n, sub.data());
    }
    if (orc_mt19937_10000th() != 4123659995u) return 3;
    std::printf("oracle: ok\n");
    return 0;
}
