// PyTorch-ROCm extension `fastergs_amd._C`: the torch-typed operations of
// the drop-in rasterizer over the C ABI of libfastergs.so
// (include/fastergs.h). It takes the role of the reference backend's
// `_C.ops` module (FasterGS gs_submodules/diff-gaussian-rasterization_faster/
// csrc/pybind.cpp:371-416): outputs and workspaces are allocated with torch
// on the inputs' device, every call runs on the current HIP stream, and ABI
// failures become RuntimeError with the library's message. The Python
// autograd function (fastergs_amd/rasterizer.py) calls these ops once per
// forward and once per backward.
//
// ops.forward: preprocess and render one frame. With a capacity hint (the
//   largest pair / bucket counts seen for frames of the same shape) the
//   render is queued at once with that capacity while the frame's true
//   counts are read on a side stream that waits for the preprocess only; a
//   frame that needed more is rendered again with its exact counts (every
//   kernel reads the true counts on the device, so outputs never depend on
//   the capacity). Without a hint the counts are read first.
//   For a FAST frame that will be differentiated it also returns the
//   backward's gradient workspace, zeroed by the render's compositing.
// ops.backward: the rasterizer adjoint and the per-Gaussian adjoint.
// ops.densify_stats: the in-place densification statistics.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fastergs.h"

namespace {

void check(int rc, const char* what) {
  if (rc != FGS_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                             "): " + fgs_last_error());
}

void* ptr(const at::Tensor& t) { return t.defined() && t.numel() > 0 ? t.data_ptr() : nullptr; }

at::Tensor f32(const at::Tensor& t, const at::Device& dev) {
  return t.defined() && t.numel() > 0 ? t.to(dev, at::kFloat).contiguous() : at::Tensor();
}

hipStream_t current_stream(const at::Device& dev) {
  return c10::hip::getCurrentHIPStream(dev.index()).stream();
}

// One default-priority non-blocking stream and event per device, for the
// count read of a speculative forward.
struct Side {
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
};
std::mutex g_side_mu;
Side g_side[64];

Side& side_for(int dev) {
  if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_side_mu);
  Side& s = g_side[dev];
  if (!s.stream) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    if (cur != dev && hipSetDevice(dev) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
    const bool ok = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) == hipSuccess &&
                    hipEventCreateWithFlags(&s.event, hipEventDisableTiming) == hipSuccess;
    if (cur != dev) (void)hipSetDevice(cur);
    if (!ok) throw std::runtime_error("side stream creation failed");
  }
  return s;
}

// The frame description shared by the ops: the caller's tensors (contiguous
// float32 on one device) and the settings.
struct Frame {
  at::Device dev{at::kCPU};
  at::Tensor means3D, opacities, scales, rotations, dc, shs, colors, culling, view, proj, campos,
      bg, mlp;
  fgs_camera cam{};
  fgs_gaussians g{};
  fgs_tile_sched ts{};
  fgs_options opt{};
};

// opts = (schedule, bucket_size, split_tiles, abs_grad, forward_only,
//         tile_mode (-1: none), split_len, features, log_cycles, order,
//         items, n_items, prev_features, features_out, composite_queue,
//         tile_w, list_prefix, list_status) --
//         items .. features_out are device addresses of the tile schedule's
//         work list and features (0: none), owned by the caller's schedule
//         object; list_status is the address of a device-accessible word
//         (0: none)
Frame make_frame(const at::Tensor& means3D, const at::Tensor& opacities, const at::Tensor& scales,
                 const at::Tensor& rotations, const at::Tensor& dc, const at::Tensor& shs,
                 const at::Tensor& colors, const at::Tensor& culling, const at::Tensor& view,
                 const at::Tensor& proj, const at::Tensor& campos, const at::Tensor& bg,
                 int64_t width, int64_t height, double tanfovx, double tanfovy,
                 double scale_modifier, double mult, int64_t sh_degree,
                 const std::vector<int64_t>& opts, const at::Tensor& mlp) {
  if (!means3D.is_cuda()) throw std::invalid_argument("means3D must be a GPU tensor");
  if (opts.size() != 18) throw std::invalid_argument("options must have 18 entries");
  Frame f;
  f.dev = means3D.device();
  const int64_t P = means3D.size(0);
  f.means3D = f32(means3D, f.dev);
  f.opacities = f32(opacities, f.dev);
  f.scales = f32(scales, f.dev);
  f.rotations = f32(rotations, f.dev);
  f.dc = f32(dc, f.dev);
  f.shs = f32(shs, f.dev);
  f.colors = f32(colors, f.dev);
  if (culling.defined() && culling.numel() > 0)
    f.culling = culling.to(f.dev, at::kBool).contiguous();
  f.view = f32(view, f.dev);
  f.proj = f32(proj, f.dev);
  f.campos = f32(campos, f.dev);
  f.bg = f32(bg, f.dev);
  f.mlp = f32(mlp, f.dev);
  const int M = f.shs.defined() ? (int)(f.shs.numel() / (3 * (P > 0 ? P : 1))) : 0;
  f.cam = fgs_camera{(int)width, (int)height, (float)tanfovx, (float)tanfovy,
                     (float)scale_modifier, (float)mult, (int)sh_degree,
                     (const float*)ptr(f.view), (const float*)ptr(f.proj),
                     (const float*)ptr(f.campos), (const float*)ptr(f.bg)};
  f.g = fgs_gaussians{(int)P, M, (const float*)ptr(f.means3D), (const float*)ptr(f.opacities),
                      (const float*)ptr(f.scales), (const float*)ptr(f.rotations),
                      (const float*)ptr(f.dc), (const float*)ptr(f.shs),
                      (const float*)ptr(f.colors), (const uint8_t*)ptr(f.culling)};
  f.opt = fgs_options{};
  f.opt.schedule = (int)opts[0];
  f.opt.bucket_size = (int)opts[1];
  f.opt.split_tiles = (int)opts[2];
  f.opt.abs_grad = (int)opts[3];
  f.opt.forward_only = (int)opts[4];
  f.opt.composite_queue = (int)opts[14];
  f.opt.tile_w = (int)opts[15];
  f.opt.list_prefix = (int)opts[16];
  f.opt.list_status = reinterpret_cast<uint32_t*>(static_cast<uintptr_t>(opts[17]));
  if (opts[5] >= 0) {
    auto addr = [](int64_t a) { return reinterpret_cast<void*>(static_cast<uintptr_t>(a)); };
    f.ts = fgs_tile_sched{(int)opts[5],
                          (int)opts[6],
                          (const float*)ptr(f.mlp),
                          (int)opts[7],
                          (int)opts[8],
                          (int)opts[9],
                          (uint32_t*)addr(opts[10]),
                          (uint32_t*)addr(opts[11]),
                          (const float*)addr(opts[12]),
                          (float*)addr(opts[13])};
    f.opt.tile_sched = &f.ts;
  }
  return f;
}

at::TensorOptions bytes(const at::Device& d) { return at::TensorOptions().dtype(at::kByte).device(d); }
at::TensorOptions floats(const at::Device& d) { return at::TensorOptions().dtype(at::kFloat).device(d); }
at::TensorOptions ints(const at::Device& d) { return at::TensorOptions().dtype(at::kInt).device(d); }

// Returns (radii, gauss_ws, pair_ws, out_color, T_final, last_contributor,
// [R, NB, R_cap, NB_cap]).
py::tuple forward(const at::Tensor& means3D, const at::Tensor& opacities, const at::Tensor& scales,
                  const at::Tensor& rotations, const at::Tensor& dc, const at::Tensor& shs,
                  const at::Tensor& colors, const at::Tensor& culling, const at::Tensor& view,
                  const at::Tensor& proj, const at::Tensor& campos, const at::Tensor& bg,
                  int64_t width, int64_t height, double tanfovx, double tanfovy,
                  double scale_modifier, double mult, int64_t sh_degree,
                  std::vector<int64_t> opts, const at::Tensor& mlp, int64_t R_hint,
                  int64_t NB_hint) {
  Frame f = make_frame(means3D, opacities, scales, rotations, dc, shs, colors, culling, view, proj,
                       campos, bg, width, height, tanfovx, tanfovy, scale_modifier, mult,
                       sh_degree, opts, mlp);
  const int P = f.g.P, W = f.cam.width, H = f.cam.height;
  hipStream_t s = current_stream(f.dev);
  at::Tensor radii = at::empty({P}, ints(f.dev));
  const size_t gn = fgs_gauss_ws_size(P, W, H, &f.opt);
  at::Tensor gws = at::empty({(int64_t)gn}, bytes(f.dev));
  check(fgs_preprocess(&f.cam, &f.g, &f.opt, gws.data_ptr(), gn, radii.data_ptr<int32_t>(), s),
        "fgs_preprocess");
  // a FAST frame that will be differentiated: its backward's gradient
  // workspace, zeroed by the render's compositing
  at::Tensor grad_ws;
  if (!f.opt.forward_only && f.opt.schedule == FGS_SCHEDULE_FAST && P > 0) {
    grad_ws = at::empty({(int64_t)fgs_grad_ws_size(P)}, bytes(f.dev));
    f.opt.zero_grad_ws = grad_ws.data_ptr();
  }
  at::Tensor out = at::empty({3, H, W}, floats(f.dev));
  at::Tensor tfinal = at::empty({H, W, 1}, floats(f.dev));
  at::Tensor contrib = at::empty({H, W, 1}, ints(f.dev));
  at::Tensor pws;
  int64_t cnt[2] = {0, 0}, cap[2] = {0, 0};
  auto render = [&]() {
    const size_t pn = fgs_pair_ws_size(P, W, H, cap[0], cap[1], &f.opt);
    pws = at::empty({(int64_t)pn}, bytes(f.dev));
    check(fgs_render(&f.cam, &f.g, &f.opt, gws.data_ptr(), gn, pws.data_ptr(), pn, cap[0], cap[1],
                     out.data_ptr<float>(), tfinal.data_ptr<float>(),
                     contrib.data_ptr<int32_t>(), s),
          "fgs_render");
  };
  if (R_hint > 0 && P > 0) {
    Side& side = side_for(f.dev.index());
    if (hipEventRecord(side.event, s) != hipSuccess ||
        hipStreamWaitEvent(side.stream, side.event, 0) != hipSuccess)
      throw std::runtime_error("stream ordering failed");
    cap[0] = R_hint;
    cap[1] = NB_hint;
    render();
    check(fgs_read_counts(gws.data_ptr(), P, W, H, &f.opt, cnt, side.stream), "fgs_read_counts");
    if (cnt[0] > cap[0] || cnt[1] > cap[1]) {
      cap[0] = cnt[0];
      cap[1] = cnt[1];
      render();
    }
  } else {
    check(fgs_read_counts(gws.data_ptr(), P, W, H, &f.opt, cnt, s), "fgs_read_counts");
    cap[0] = cnt[0];
    cap[1] = cnt[1];
    render();
  }
  return py::make_tuple(radii, gws, pws, out, tfinal, contrib,
                        std::vector<int64_t>{cnt[0], cnt[1], cap[0], cap[1]}, grad_ws);
}

// Returns (dmean2D, dconic, dopacity, dcolor, dmean3D, dcov3D, ddc, dsh,
// dscale, drot, grad_ws); dconic, dcolor, dcov3D are undefined unless wanted.
// zeroed_grad_ws: the gradient workspace ops.forward returned (zeroed by the
// frame's render), passed to one backward of the frame, or empty.
// stats: (xyz_gradient_accum, xyz_gradient_accum_abs, max_radii, denom),
// each updated in place when defined and non-empty (else left alone).
py::tuple backward(const at::Tensor& means3D, const at::Tensor& opacities, const at::Tensor& scales,
                   const at::Tensor& rotations, const at::Tensor& dc, const at::Tensor& shs,
                   const at::Tensor& colors, const at::Tensor& culling, const at::Tensor& view,
                   const at::Tensor& proj, const at::Tensor& campos, const at::Tensor& bg,
                   int64_t width, int64_t height, double tanfovx, double tanfovy,
                   double scale_modifier, double mult, int64_t sh_degree,
                   std::vector<int64_t> opts, const at::Tensor& mlp, const at::Tensor& gws,
                   const at::Tensor& pws, int64_t R_cap, int64_t NB_cap, const at::Tensor& radii,
                   const at::Tensor& tfinal, const at::Tensor& contrib,
                   const at::Tensor& grad_out, bool want_conic, bool want_colors,
                   bool want_cov3D, const std::vector<at::Tensor>& stats,
                   const at::Tensor& zeroed_grad_ws) {
  Frame f = make_frame(means3D, opacities, scales, rotations, dc, shs, colors, culling, view, proj,
                       campos, bg, width, height, tanfovx, tanfovy, scale_modifier, mult,
                       sh_degree, opts, mlp);
  const int64_t P = f.g.P, M = f.g.max_coeffs;
  hipStream_t s = current_stream(f.dev);
  at::Tensor g = f32(grad_out, f.dev);
  if (g.numel() != 3 * (int64_t)f.cam.width * f.cam.height)
    throw std::invalid_argument("grad_out must hold 3 x H x W values");
  auto e = [&](std::vector<int64_t> shape) { return at::empty(shape, floats(f.dev)); };
  at::Tensor dmean2D = e({P, 4}), dopac = e({P, 1}), dmean3D = e({P, 3}), ddc = e({P, 1, 3}),
             dsh = e({P, M, 3}), dscale = e({P, 3}), drot = e({P, 4});
  at::Tensor dconic = want_conic ? e({P, 2, 2}) : at::Tensor();
  at::Tensor dcolor = want_colors ? e({P, 3}) : at::Tensor();
  at::Tensor dcov = want_cov3D ? e({P, 6}) : at::Tensor();
  // the forward's gradient workspace, zeroed by its preprocess and not used
  // since (the caller passes it once), else a new one the backward zeroes
  const size_t gw_bytes = fgs_grad_ws_size((int)P);
  at::Tensor grad_ws;
  if (zeroed_grad_ws.defined() && zeroed_grad_ws.is_cuda() &&
      zeroed_grad_ws.device() == f.dev && zeroed_grad_ws.is_contiguous() &&
      (size_t)zeroed_grad_ws.numel() * zeroed_grad_ws.element_size() >= gw_bytes) {
    grad_ws = zeroed_grad_ws;
    f.opt.grad_ws_zeroed = 1;
  } else {
    grad_ws = at::empty({(int64_t)gw_bytes}, bytes(f.dev));
  }
  if (stats.size() != 4) throw std::invalid_argument("stats must have 4 entries");
  float* st[4];
  for (int k = 0; k < 4; ++k) {
    const at::Tensor& t = stats[k];
    st[k] = nullptr;
    if (!t.defined() || t.numel() == 0) continue;
    if (!t.is_cuda() || t.scalar_type() != at::kFloat || !t.is_contiguous() || t.numel() != P)
      throw std::invalid_argument("densification statistics must be contiguous float32 GPU "
                                  "tensors of P entries");
    st[k] = t.data_ptr<float>();
  }
  const fgs_grads out{(float*)ptr(dmean2D), (float*)ptr(dconic), (float*)ptr(dopac),
                      (float*)ptr(dcolor),  (float*)ptr(dmean3D), (float*)ptr(dcov),
                      (float*)ptr(ddc),     (float*)ptr(dsh),     (float*)ptr(dscale),
                      (float*)ptr(drot),    st[0],                st[1],
                      st[2],                st[3]};
  check(fgs_backward(&f.cam, &f.g, &f.opt, ptr(gws), ptr(pws), R_cap, NB_cap,
                     (const int32_t*)ptr(radii), (const float*)ptr(tfinal),
                     (const int32_t*)ptr(contrib), g.data_ptr<float>(), grad_ws.data_ptr(), &out,
                     s),
        "fgs_backward");
  return py::make_tuple(dmean2D, dconic, dopac, dcolor, dmean3D, dcov, ddc, dsh, dscale, drot,
                        grad_ws);
}

void densify_stats(const at::Tensor& radii, const at::Tensor& dmean2D, const at::Tensor& accum,
                   const at::Tensor& accum_abs, const at::Tensor& max_radii,
                   const at::Tensor& denom) {
  const int P = (int)radii.numel();
  for (const at::Tensor* t : {&accum, &accum_abs, &max_radii, &denom})
    if (t->defined() && t->numel() > 0 &&
        (t->scalar_type() != at::kFloat || !t->is_contiguous() || t->numel() != P))
      throw std::invalid_argument("statistics must be contiguous float32 with one value per Gaussian");
  check(fgs_densify_stats(P, (const int32_t*)ptr(radii), (const float*)ptr(dmean2D),
                          (float*)ptr(accum), (float*)ptr(accum_abs), (float*)ptr(max_radii),
                          (float*)ptr(denom), current_stream(radii.device())),
        "fgs_densify_stats");
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "torch operations of the gfx950 Gaussian rasterizer (libfastergs.so)";
  py::module_ ops = m.def_submodule("ops", "rasterizer operations");
  ops.def("forward", &forward, "preprocess + render one frame");
  ops.def("backward", &backward, "rasterizer and per-Gaussian adjoint of one frame");
  ops.def("densify_stats", &densify_stats, "in-place densification statistics");
  m.def("build_info", []() { return std::string(fgs_build_info()); });
}
