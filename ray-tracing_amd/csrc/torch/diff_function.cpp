// torch::autograd::Function over the differentiable renderer's C ABI: the
// forward is lrt_diff_render (one path-traced image of theta), the backward is
// lrt_diff_adjoint (path replay of other samples than the forward's, so a loss
// gradient computed from the forward image does not correlate with the
// replayed paths).  Float32 tensors on the scene's GPU; theta is the flat
// parameter vector (texels, two GGX alphas, one IOR), the image H x W x 3.
//
// The library's entry points are looked up in the already loaded
// liblinrt_hip.so (bind), so this module and the Python package share one
// copy of the library and its scene handles.
#include <dlfcn.h>

#include <string>
#include <vector>

#include <torch/cuda.h>
#include <torch/extension.h>

#include "linrt.h"

namespace {

struct Api {
    decltype(&lrt_diff_render) render = nullptr;
    decltype(&lrt_diff_adjoint) adjoint = nullptr;
    decltype(&lrt_diff_param_count) count = nullptr;
    decltype(&lrt_last_error) last_error = nullptr;
} api;

template <class F>
void lookup(void* h, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(h, name));
    TORCH_CHECK(f != nullptr, "liblinrt_hip.so has no symbol ", name);
}

// path: the liblinrt_hip.so the Python package loaded
void bind(const std::string& path) {
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_NOLOAD);
    TORCH_CHECK(h != nullptr, "liblinrt_hip.so is not loaded in this process: ", path);
    lookup(h, "lrt_diff_render", api.render);
    lookup(h, "lrt_diff_adjoint", api.adjoint);
    lookup(h, "lrt_diff_param_count", api.count);
    lookup(h, "lrt_last_error", api.last_error);
}

void check(int rc, const char* what) { TORCH_CHECK(rc == LRT_OK, what, " failed (", rc, "): ", api.last_error()); }

struct Job {
    lrt_scene* scene;
    lrt_diff_params dp;
    lrt_render_params prm;
};

// dp: tex_material, tex_width, tex_height, alpha_material[0], alpha_material[1],
// ior_material; prm: the ten fields of lrt_render_params in order
Job make_job(int64_t scene, const std::vector<int64_t>& dp, const std::vector<int64_t>& prm) {
    TORCH_CHECK(api.render != nullptr, "linrt_torch: bind() the library first");
    TORCH_CHECK(scene != 0, "null scene handle");
    TORCH_CHECK(dp.size() == 6, "dp: 6 integers expected, got ", dp.size());
    TORCH_CHECK(prm.size() == 10, "prm: 10 integers expected, got ", prm.size());
    Job j;
    j.scene = reinterpret_cast<lrt_scene*>(scene);
    j.dp.tex_material = (int32_t)dp[0];
    j.dp.tex_width = (uint32_t)dp[1];
    j.dp.tex_height = (uint32_t)dp[2];
    j.dp.alpha_material[0] = (int32_t)dp[3];
    j.dp.alpha_material[1] = (int32_t)dp[4];
    j.dp.ior_material = (int32_t)dp[5];
    j.prm.width = (uint32_t)prm[0];
    j.prm.height = (uint32_t)prm[1];
    j.prm.spp = (uint32_t)prm[2];
    j.prm.spp_offset = (uint32_t)prm[3];
    j.prm.max_depth = (uint32_t)prm[4];
    j.prm.rr_depth = (uint32_t)prm[5];
    j.prm.seed = (uint64_t)prm[6];
    j.prm.row_begin = (uint32_t)prm[7];
    j.prm.row_end = (uint32_t)prm[8];
    j.prm.flags = (uint32_t)prm[9];
    return j;
}

void check_theta(const torch::Tensor& theta, const Job& j) {
    TORCH_CHECK(theta.is_cuda(), "theta must be a GPU tensor");
    TORCH_CHECK(theta.scalar_type() == torch::kFloat && theta.dim() == 1, "theta must be a 1-d float32 tensor");
    uint32_t n = 0;
    check(api.count(&j.dp, &n), "lrt_diff_param_count");
    TORCH_CHECK(theta.numel() == (int64_t)n, "theta has ", theta.numel(), " entries, the parameters are ", n);
}

class DiffRender : public torch::autograd::Function<DiffRender> {
   public:
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, const torch::Tensor& theta, int64_t scene,
                                 std::vector<int64_t> dp, std::vector<int64_t> prm) {
        const Job j = make_job(scene, dp, prm);
        const torch::Tensor th = theta.detach().contiguous();
        check_theta(th, j);
        torch::Tensor image = torch::empty({(int64_t)j.prm.height, (int64_t)j.prm.width, 3}, th.options());
        // the library works on its own stream: theta must be written first
        torch::cuda::synchronize(th.get_device());
        check(api.render(j.scene, &j.dp, th.data_ptr<float>(), &j.prm, image.data_ptr<float>(), 1), "lrt_diff_render");
        ctx->save_for_backward({th});
        ctx->saved_data["scene"] = scene;
        ctx->saved_data["dp"] = dp;
        ctx->saved_data["prm"] = prm;
        return image;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out) {
        Job j = make_job(ctx->saved_data["scene"].toInt(), ctx->saved_data["dp"].toIntVector(),
                         ctx->saved_data["prm"].toIntVector());
        j.prm.spp_offset += 1u << 30;  // other samples than the forward image's
        const torch::Tensor th = ctx->get_saved_variables()[0];
        const torch::Tensor d_img = grad_out[0].to(torch::kFloat).contiguous();
        TORCH_CHECK(d_img.dim() == 3 && d_img.size(0) == (int64_t)j.prm.height && d_img.size(1) == (int64_t)j.prm.width &&
                        d_img.size(2) == 3,
                    "the image gradient must be H x W x 3");
        torch::Tensor grad = torch::empty_like(th);
        torch::cuda::synchronize(th.get_device());
        check(api.adjoint(j.scene, &j.dp, th.data_ptr<float>(), &j.prm, d_img.data_ptr<float>(), grad.data_ptr<float>(), 1),
              "lrt_diff_adjoint");
        return {grad, torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor render(const torch::Tensor& theta, int64_t scene, std::vector<int64_t> dp, std::vector<int64_t> prm) {
    return DiffRender::apply(theta, scene, std::move(dp), std::move(prm));
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "torch::autograd::Function over lrt_diff_render / lrt_diff_adjoint";
    m.def("bind", &bind, "look up the C ABI in the loaded liblinrt_hip.so at this path");
    m.def("render", &render,
          "image = render(theta, scene handle, dp[6], prm[10]); differentiable in theta by path replay");
}
