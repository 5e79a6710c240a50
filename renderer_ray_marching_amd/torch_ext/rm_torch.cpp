// libtorch front end of the differentiable march: a torch::autograd::Function whose forward
// renders all views in one batched march (rm_batch_render_diff: hit selection by the march,
// analytic sphere depth, flat point colour) and whose backward runs the adjoint kernels
// (rm_render_adjoint per view), which reduce per-point gradients by sorting the hit pixels by
// point index. All arithmetic runs in librm_hip.so; this file only moves tensors across the
// C ABI.
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
// the scene therefore owns).
struct DiffScene : torch::CustomClassHolder {
    rm_ctx* ctx = nullptr;
    rm_cloud* cloud = nullptr;
    int64_t n = 0;
    std::vector<ViewSpec> views;
    rm_batch* batch = nullptr;
    torch::Tensor rgba_buf, hit_buf;

    DiffScene(int64_t ctx_ptr, int64_t cloud_ptr, int64_t n_points)
        : ctx(reinterpret_cast<rm_ctx*>(ctx_ptr)),
          cloud(reinterpret_cast<rm_cloud*>(cloud_ptr)),
          n(n_points) {}

    ~DiffScene() override { drop_batch(); }

    void drop_batch() {
        if (batch) rm_batch_destroy(batch);
        batch = nullptr;
    }

    void add_view(std::vector<double> pose, double fov, int64_t W, int64_t H) {
        TORCH_CHECK(pose.size() == 16, "pose must have 16 entries");
        ViewSpec v;
        v.pose.assign(pose.begin(), pose.end());
        v.fov = (float)fov;
        v.W = (int)W;
        v.H = (int)H;
        views.push_back(v);
        drop_batch();
    }

    void ensure_batch(const torch::TensorOptions& fopt) {
        if (batch) return;
        const int64_t V = (int64_t)views.size();
        const int64_t W = views[0].W, H = views[0].H;
        rgba_buf = torch::empty({V * H * W * 4}, fopt);
        hit_buf = torch::empty({V, H, W}, fopt.dtype(torch::kInt32));
        std::vector<rm_view> rv((size_t)V);
        for (int64_t v = 0; v < V; ++v) {
            std::copy(views[v].pose.begin(), views[v].pose.end(), rv[v].pose);
            rv[v].fov_radians = views[v].fov;
            rv[v].width = (int32_t)W;
            rv[v].height = (int32_t)H;
            rv[v].reserved = 0;
            rv[v].offset = v * H * W;
        }
        check_rm(rm_batch_create(ctx, cloud, rv.data(), (int)V, nullptr,
                                 rgba_buf.data_ptr<float>(), hit_buf.data_ptr<int32_t>(), &batch),
                 ctx, "rm_batch_create");
    }
};

void bind_params(DiffScene& s, const torch::Tensor& xyz, const torch::Tensor& radii,
                 const torch::Tensor& rgb) {
    TORCH_CHECK(xyz.is_cuda() && xyz.dtype() == torch::kFloat32 && xyz.dim() == 2 &&
                    xyz.size(0) == s.n && xyz.size(1) == 3,
                "xyz must be a (n, 3) float32 CUDA tensor");
    TORCH_CHECK(radii.is_cuda() && radii.dtype() == torch::kFloat32 && radii.numel() == s.n,
                "radii must be n float32 on the GPU");
    TORCH_CHECK(rgb.is_cuda() && rgb.dtype() == torch::kFloat32 && rgb.dim() == 2 &&
                    rgb.size(0) == s.n && rgb.size(1) == 3,
                "rgb must be a (n, 3) float32 CUDA tensor");
    auto x = xyz.contiguous();
    auto r = radii.contiguous();
    auto c4 = torch::cat({rgb, torch::ones({s.n, 1}, rgb.options())}, 1).contiguous();
    check_rm(rm_cloud_refit(s.ctx, s.cloud, x.data_ptr<float>()), s.ctx, "rm_cloud_refit");
    check_rm(rm_cloud_set_radii(s.ctx, s.cloud, r.data_ptr<float>()), s.ctx, "rm_cloud_set_radii");
    check_rm(rm_cloud_set_colors(s.ctx, s.cloud, c4.data_ptr<float>()), s.ctx,
             "rm_cloud_set_colors");
}

class RenderFunction : public torch::autograd::Function<RenderFunction> {
   public:
    static torch::autograd::variable_list forward(torch::autograd::AutogradContext* actx,
                                                  torch::Tensor xyz, torch::Tensor radii,
                                                  torch::Tensor rgb,
                                                  c10::intrusive_ptr<DiffScene> scene) {
        DiffScene& s = *scene;
        bind_params(s, xyz, radii, rgb);
        int64_t V = (int64_t)s.views.size();
        TORCH_CHECK(V > 0, "no views");
        int W = s.views[0].W, H = s.views[0].H;
        for (auto& v : s.views) TORCH_CHECK(v.W == W && v.H == H, "views must share a size");
        auto fopt = xyz.options();
        auto img = torch::empty({V, H, W, 3}, fopt);
        auto depth = torch::empty({V, H, W}, fopt);
        s.ensure_batch(fopt);
        check_rm(rm_batch_render_diff(s.batch, img.data_ptr<float>(), depth.data_ptr<float>()),
                 s.ctx, "rm_batch_render_diff");
        // the batch reuses its hit buffer at the next forward; backward needs this one
        auto hit = s.hit_buf.clone();
        actx->saved_data["scene"] = scene;
        actx->save_for_backward({hit});
        actx->mark_non_differentiable({hit});
        return {img, depth, hit};
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* actx,
                                                   torch::autograd::variable_list grads) {
        auto scene = actx->saved_data["scene"].toCustomClass<DiffScene>();
        DiffScene& s = *scene;
        auto hit = actx->get_saved_variables()[0];
        int64_t V = hit.size(0), H = hit.size(1), W = hit.size(2);
        auto fopt = hit.options().dtype(torch::kFloat32);
        auto g_img = grads[0].defined() ? grads[0].contiguous() : torch::zeros({V, H, W, 3}, fopt);
        auto g_depth = grads[1].defined() ? grads[1].contiguous() : torch::zeros({V, H, W}, fopt);
        auto g_xyz = torch::zeros({s.n, 3}, fopt);
        auto g_r = torch::zeros({s.n}, fopt);
        auto g_rgb = torch::zeros({s.n, 3}, fopt);
        for (int64_t v = 0; v < V; ++v) {
            auto& vs = s.views[v];
            check_rm(rm_render_adjoint(s.ctx, s.cloud, vs.pose.data(), vs.fov, (int)W, (int)H,
                                       hit[v].data_ptr<int32_t>(), g_img[v].data_ptr<float>(),
                                       g_depth[v].data_ptr<float>(), g_xyz.data_ptr<float>(),
                                       g_r.data_ptr<float>(), g_rgb.data_ptr<float>()),
                     s.ctx, "rm_render_adjoint");
        }
        return {g_xyz, g_r, g_rgb, torch::Tensor()};
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
        .def("add_view", &DiffScene::add_view);
    m.def("render", &render);
}
