// dol_torch.cpp -- libtorch autograd Function over the continuous-seed
// distance of libdolhip (dolhip_seed_nearest_2d_dev / dolhip_seed_grad_2d_dev).
//
//   dist, idx = seed_distance(seeds, w, h)
//
// seeds: (n, 2) float32 on the GPU, cell centres at integer (x, y).  dist is
// differentiable in seeds; idx (int32, nearest seed per cell) is not.  The
// backward pass reuses the forward's exact assignment and distances.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include "dol_hip.h"

namespace {

void check(int rc, const char* what)
{
    TORCH_CHECK(rc == DOLHIP_OK, what, " failed: ", dolhip_last_error());
}

void* current_stream() { return (void*)c10::hip::getCurrentHIPStream().stream(); }

class SeedDistance : public torch::autograd::Function<SeedDistance> {
public:
    static torch::autograd::variable_list forward(torch::autograd::AutogradContext* ctx,
                                                  torch::Tensor seeds, int64_t w, int64_t h)
    {
        TORCH_CHECK(seeds.is_cuda() && seeds.scalar_type() == torch::kFloat32 && seeds.dim() == 2 &&
                        seeds.size(1) == 2,
                    "seeds must be a float32 (n, 2) GPU tensor");
        auto s = seeds.contiguous();
        const int n = (int)s.size(0);
        auto opts = s.options();
        auto dist = torch::empty({h, w}, opts);
        auto idx = torch::empty({h, w}, opts.dtype(torch::kInt32));
        auto work = torch::empty({(int64_t)dolhip_seed_workspace((int)w, (int)h, n)}, opts.dtype(torch::kUInt8));
        check(dolhip_seed_nearest_2d_dev((int)w, (int)h, n, s.data_ptr<float>(), idx.data_ptr<int32_t>(),
                                         dist.data_ptr<float>(), work.data_ptr(), current_stream()),
              "dolhip_seed_nearest_2d_dev");
        ctx->save_for_backward({s, idx, dist});
        ctx->saved_data["w"] = w;
        ctx->saved_data["h"] = h;
        ctx->mark_non_differentiable({idx});
        return {dist, idx};
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grads)
    {
        auto saved = ctx->get_saved_variables();
        auto s = saved[0], idx = saved[1], dist = saved[2];
        const int w = (int)ctx->saved_data["w"].toInt(), h = (int)ctx->saved_data["h"].toInt();
        auto gout = grads[0].contiguous().to(torch::kFloat32);
        auto grad = torch::empty_like(s);
        check(dolhip_seed_grad_2d_dev(w, h, (int)s.size(0), s.data_ptr<float>(), idx.data_ptr<int32_t>(),
                                      dist.data_ptr<float>(), gout.data_ptr<float>(), grad.data_ptr<float>(),
                                      current_stream()),
              "dolhip_seed_grad_2d_dev");
        return {grad, torch::Tensor(), torch::Tensor()};
    }
};

std::vector<torch::Tensor> seed_distance(torch::Tensor seeds, int64_t w, int64_t h)
{
    auto r = SeedDistance::apply(seeds, w, h);
    return {r[0], r[1]};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.def("seed_distance", &seed_distance,
          "distance to the nearest continuous seed and its index; differentiable in the seeds");
}
