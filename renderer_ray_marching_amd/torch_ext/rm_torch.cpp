// libtorch front end of the differentiable march: a torch::autograd::Function whose forward
// renders all views in one batched march (rm_batch_render_diff: hit selection by the march,
// analytic sphere depth, flat point colour) and whose backward runs the batch adjoint
// (rm_batch_render_adjoint: one sort of every view's hit pixels by point, one reduction per
// point) at the centres and radii the forward marched. All arithmetic runs in librm_hip.so;
// this file only moves tensors across the C ABI. Both directions run on the caller's current
// stream (the context is pointed at it for the call and back at its own stream after it), so
// they are ordered with the torch work that produces their inputs and consumes their outputs.
#include <c10/hip/HIPStream.h>
#include <torch/script.h>
#include <torch/torch.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/rm.h"

namespace {

void check_rm(int code, rm_ctx* ctx, const char* what) {
    TORCH_CHECK(code == RM_OK, what, " failed (", code, "): ", rm_last_error(ctx));
}

struct ViewSpec {
    std::vector<float> pose;  // 16, column-major
    float fov;
    int W, H;
};

// Scene state shared by forward and backward: the library context, the tree built over the
// initial points (refitted to every new set of centres), the cloud bound to it, and a batch over
// all views (created at the first forward; its graph captures the march's output buffers, which
// the scene therefore owns). Views may differ in size: their pixels are concatenated view after
// view.
struct DiffScene : torch::CustomClassHolder {
    rm_ctx* ctx = nullptr;
    rm_cloud* cloud = nullptr;
    int64_t n = 0;
    std::vector<ViewSpec> views;
    rm_batch* batch = nullptr;
    int64_t generation = 0;  // bumped whenever the batch is dropped (views changed)
    torch::Tensor rgba_buf, hit_buf;

    DiffScene(int64_t ctx_ptr, int64_t cloud_ptr, int64_t n_points)
        : ctx(reinterpret_cast<rm_ctx*>(ctx_ptr)),
          cloud(reinterpret_cast<rm_cloud*>(cloud_ptr)),
          n(n_points) {}

    ~DiffScene() override { drop_batch(); }

    void drop_batch() {
        if (batch) rm_batch_destroy(batch);
        batch = nullptr;
        ++generation;
    }

    void add_view(std::vector<double> pose, double fov, int64_t W, int64_t H) {
        TORCH_CHECK(pose.size() == 16, "pose must have 16 entries");
        TORCH_CHECK(W > 0 && H > 0, "empty view");
        ViewSpec v;
        v.pose.assign(pose.begin(), pose.end());
        v.fov = (float)fov;
        v.W = (int)W;
        v.H = (int)H;
        views.push_back(v);
        drop_batch();
    }

    // First pixel of each view in the concatenated outputs, then the total.
    std::vector<int64_t> view_offsets() const {
        std::vector<int64_t> off(views.size() + 1, 0);
        for (size_t v = 0; v < views.size(); ++v)
            off[v + 1] = off[v] + (int64_t)views[v].W * views[v].H;
        return off;
    }

    bool same_size() const {
        for (auto& v : views)
            if (v.W != views[0].W || v.H != views[0].H) return false;
        return true;
    }

    void ensure_batch(const torch::TensorOptions& fopt) {
        if (batch) return;
        const auto off = view_offsets();
        const int64_t V = (int64_t)views.size(), P = off.back();
        rgba_buf = torch::empty({P * 4}, fopt);
        hit_buf = torch::empty({P}, fopt.dtype(torch::kInt32));
        std::vector<rm_view> rv((size_t)V);
        for (int64_t v = 0; v < V; ++v) {
            std::copy(views[v].pose.begin(), views[v].pose.end(), rv[v].pose);
            rv[v].fov_radians = views[v].fov;
            rv[v].width = (int32_t)views[v].W;
            rv[v].height = (int32_t)views[v].H;
            rv[v].reserved = 0;
            rv[v].offset = off[v];
        }
        check_rm(rm_batch_create(ctx, cloud, rv.data(), (int)V, nullptr,
                                 rgba_buf.data_ptr<float>(), hit_buf.data_ptr<int32_t>(), &batch),
                 ctx, "rm_batch_create");
    }
};

// Points the library context at torch's current stream on the tensors' device for the lifetime
// of the guard, then back at the stream it had.
struct CurrentStream {
    rm_ctx* ctx;
    void* prev = nullptr;
    CurrentStream(rm_ctx* c, const torch::Tensor& t) : ctx(c) {
        check_rm(rm_ctx_get_stream(ctx, &prev), ctx, "rm_ctx_get_stream");
        hipStream_t st = c10::hip::getCurrentHIPStream(t.device().index()).stream();
        check_rm(rm_ctx_set_stream(ctx, reinterpret_cast<void*>(st)), ctx, "rm_ctx_set_stream");
    }
    ~CurrentStream() { (void)rm_ctx_set_stream(ctx, prev); }
    CurrentStream(const CurrentStream&) = delete;
    CurrentStream& operator=(const CurrentStream&) = delete;
};

void bind_params(DiffScene& s, const torch::Tensor& x, const torch::Tensor& r,
                 const torch::Tensor& rgb) {
    auto c4 = torch::cat({rgb, torch::ones({s.n, 1}, rgb.options())}, 1).contiguous();
    check_rm(rm_cloud_refit(s.ctx, s.cloud, x.data_ptr<float>()), s.ctx, "rm_cloud_refit");
    check_rm(rm_cloud_set_radii(s.ctx, s.cloud, r.data_ptr<float>()), s.ctx, "rm_cloud_set_radii");
    check_rm(rm_cloud_set_colors(s.ctx, s.cloud, c4.data_ptr<float>()), s.ctx,
             "rm_cloud_set_colors");
}

class RenderFunction : public torch::autograd::Function<RenderFunction> {
   public:
    // Same-size views: img (V, H, W, 3), depth (V, H, W), hit (V, H, W). Mixed sizes: the views
    // concatenated, img (P, 3), depth (P), hit (P), view v at scene.view_offsets()[v].
    static torch::autograd::variable_list forward(torch::autograd::AutogradContext* actx,
                                                  torch::Tensor xyz, torch::Tensor radii,
                                                  torch::Tensor rgb,
                                                  c10::intrusive_ptr<DiffScene> scene) {
        DiffScene& s = *scene;
        TORCH_CHECK(xyz.is_cuda() && xyz.dtype() == torch::kFloat32 && xyz.dim() == 2 &&
                        xyz.size(0) == s.n && xyz.size(1) == 3,
                    "xyz must be a (n, 3) float32 CUDA tensor");
        TORCH_CHECK(radii.is_cuda() && radii.dtype() == torch::kFloat32 && radii.numel() == s.n,
                    "radii must be n float32 on the GPU");
        TORCH_CHECK(rgb.is_cuda() && rgb.dtype() == torch::kFloat32 && rgb.dim() == 2 &&
                        rgb.size(0) == s.n && rgb.size(1) == 3,
                    "rgb must be a (n, 3) float32 CUDA tensor");
        const int64_t V = (int64_t)s.views.size();
        TORCH_CHECK(V > 0, "no views");
        // the centres and radii this forward marches, kept for its backward (a later forward
        // moves the cloud; in-place edits of these tensors make autograd raise)
        auto x = xyz.contiguous();
        auto r = radii.contiguous().reshape({s.n});
        CurrentStream on_stream(s.ctx, x);
        bind_params(s, x, r, rgb);
        auto fopt = xyz.options();
        const int64_t P = s.view_offsets().back();
        auto img = torch::empty({P, 3}, fopt);
        auto depth = torch::empty({P}, fopt);
        s.ensure_batch(fopt);
        check_rm(rm_batch_render_diff(s.batch, img.data_ptr<float>(), depth.data_ptr<float>()),
                 s.ctx, "rm_batch_render_diff");
        // the batch reuses its hit buffer at the next forward; backward needs this one
        auto hit = s.hit_buf.clone();
        if (s.same_size()) {
            const int64_t H = s.views[0].H, W = s.views[0].W;
            img = img.view({V, H, W, 3});
            depth = depth.view({V, H, W});
            hit = hit.view({V, H, W});
        }
        actx->saved_data["scene"] = scene;
        actx->saved_data["generation"] = s.generation;
        actx->saved_data["radii_sizes"] = radii.sizes().vec();
        actx->save_for_backward({hit, x, r});
        actx->mark_non_differentiable({hit});
        return {img, depth, hit};
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* actx,
                                                   torch::autograd::variable_list grads) {
        auto scene = actx->saved_data["scene"].toCustomClass<DiffScene>();
        DiffScene& s = *scene;
        TORCH_CHECK(s.batch && actx->saved_data["generation"].toInt() == s.generation,
                    "the scene's views changed between forward and backward");
        auto saved = actx->get_saved_variables();
        auto hit = saved[0].reshape({-1});
        auto x = saved[1], r = saved[2];
        CurrentStream on_stream(s.ctx, x);
        const int64_t P = hit.numel();
        auto fopt = hit.options().dtype(torch::kFloat32);
        auto g_img = grads[0].defined() ? grads[0].reshape({P, 3}).contiguous()
                                        : torch::zeros({P, 3}, fopt);
        auto g_depth = grads[1].defined() ? grads[1].reshape({P}).contiguous()
                                          : torch::zeros({P}, fopt);
        auto g_xyz = torch::zeros({s.n, 3}, fopt);
        auto g_r = torch::zeros({s.n}, fopt);
        auto g_rgb = torch::zeros({s.n, 3}, fopt);
        check_rm(rm_batch_render_adjoint(s.batch, x.data_ptr<float>(), r.data_ptr<float>(),
                                         hit.data_ptr<int32_t>(), g_img.data_ptr<float>(),
                                         g_depth.data_ptr<float>(), g_xyz.data_ptr<float>(),
                                         g_r.data_ptr<float>(), g_rgb.data_ptr<float>()),
                 s.ctx, "rm_batch_render_adjoint");
        return {g_xyz, g_r.view(actx->saved_data["radii_sizes"].toIntVector()), g_rgb,
                torch::Tensor()};
    }
};

std::vector<torch::Tensor> render(torch::Tensor xyz, torch::Tensor radii, torch::Tensor rgb,
                                  c10::intrusive_ptr<DiffScene> scene) {
    return RenderFunction::apply(xyz, radii, rgb, scene);
}

}  // namespace

TORCH_LIBRARY(rm_diff, m) {
    m.class_<DiffScene>("DiffScene")
        .def(torch::init<int64_t, int64_t, int64_t>())
        .def("add_view", &DiffScene::add_view)
        .def("view_offsets", &DiffScene::view_offsets);
    m.def("render", &render);
}
