// fg_glue.cpp -- torch extension "_fastgeodis_hip" over the gfx950 C ABI.
//
// Mirrors the reference's binding module (FastGeodis/fastgeodis.cpp:44-146,
// 241-258, 319-343, 423-451) for the raster-scan entry points: the same
// function names, argument order and meaning, the same validation order and
// exception types (common.h:53-120: std::invalid_argument -> ValueError,
// std::runtime_error -> RuntimeError).  The compute always runs on the GPU:
// HIP tensors go straight to the C ABI on the current torch stream; CPU
// tensors (which the reference computes with OpenMP) are staged through the
// current HIP device and the result is returned on the CPU.  Without a GPU a
// call with CPU tensors raises RuntimeError: there is no CPU implementation.
#include <torch/extension.h>

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>

#include <cstdlib>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "fastgeodis_hip.h"

namespace {

// ---- validation (same checks, order and messages as the reference's common.h)

void print_shape(const torch::Tensor &data) {
  std::cout << "Shape: (";
  for (int64_t d = 0; d < data.dim(); ++d) {
    std::cout << data.size(d);
    std::cout << (d != data.dim() - 1 ? ", " : ")\n");
  }
}

void check_spatial_shape_match(const torch::Tensor &in1, const torch::Tensor &in2, int dims) {
  if (in1.dim() != in2.dim())
    throw std::invalid_argument("dimensions of input tensors do not match " +
                                std::to_string(in1.dim() - 2) + " vs " + std::to_string(in2.dim() - 2));
  for (int i = 0; i < dims; ++i) {
    if (in1.size(2 + i) != in2.size(2 + i)) {
      std::cout << "Tensor1 ";
      print_shape(in1);
      std::cout << "Tensor2 ";
      print_shape(in2);
      throw std::invalid_argument("shapes of input tensors do not match");
    }
  }
}

void check_cpu(const torch::Tensor &in) {
  if (in.is_cuda()) throw std::invalid_argument("input is not on CPU device, try using data.to('cpu') on input");
}

void check_cuda(const torch::Tensor &in) {
  if (!in.is_cuda()) throw std::invalid_argument("input is not on CUDA device, try using data.to('cuda') on input");
}

void check_single_batch(const torch::Tensor &in) {
  if (in.size(0) != 1) throw std::invalid_argument("FastGeodis currently only supports single batch input.");
}

void check_data_dim(const torch::Tensor &in, int dims) {
  const int num_dims = (int)in.dim();
  if (num_dims != dims)
    throw std::invalid_argument("function only supports " + std::to_string(dims - 2) +
                                "D spatial inputs, received " + std::to_string(num_dims - 2));
}

void check_input_dimensions(const torch::Tensor &image, const torch::Tensor &mask, int num_dims) {
  check_data_dim(image, num_dims);
  check_data_dim(mask, num_dims);
  check_single_batch(image);
  check_single_batch(mask);
  check_spatial_shape_match(image, mask, num_dims - 2);
}

// The reference reads both tensors through accessor<float> / data_ptr<float>,
// which raise RuntimeError for other dtypes.
void check_float(const torch::Tensor &t, const char *name) {
  if (t.scalar_type() != torch::kFloat32)
    throw std::runtime_error(std::string("expected scalar type Float but found ") +
                             c10::toString(t.scalar_type()) + " for " + name);
}

void raise_for(int rc) {
  if (rc == FG_OK) return;
  const std::string msg = fg_last_error();
  if (rc == FG_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error("fastgeodis_amd: " + msg);
}

bool sync_check_enabled() {
  const char *e = std::getenv("FASTGEODIS_SYNC_CHECK");
  return e && e[0] && e[0] != '0';
}

// ---- one call on the GPU --------------------------------------------------

struct Call {
  int op;       // FG_OP_*
  int nd;       // 2 or 3 spatial dims
  std::vector<float> spacing;
  float v, lg, le, theta;
  int iterations;
};

torch::Tensor run_device(const Call &c, const torch::Tensor &image_in, const torch::Tensor &mask_in) {
  check_float(image_in, "image");
  check_float(mask_in, "mask");
  if (image_in.device() != mask_in.device())
    throw std::invalid_argument("image and mask are on different devices");
  c10::DeviceGuard guard(image_in.device());
  const torch::Tensor image = image_in.contiguous();
  const torch::Tensor mask = mask_in.contiguous();
  torch::Tensor out = torch::empty_like(mask, torch::MemoryFormat::Contiguous);
  const int C = (int)image.size(1), Cm = (int)mask.size(1);
  int D = 1, H, W;
  if (c.nd == 2) {
    H = (int)image.size(2);
    W = (int)image.size(3);
  } else {
    D = (int)image.size(2);
    H = (int)image.size(3);
    W = (int)image.size(4);
  }
  if (out.numel() == 0) return out;
  if (C == 0) throw std::invalid_argument("image must have at least one channel");
  const size_t wsb = fg_workspace_bytes(c.op, C, c.nd == 2 ? 0 : D, H, W);
  torch::Tensor ws = torch::empty({(int64_t)wsb}, image.options().dtype(torch::kUInt8));
  void *stream = (void *)c10::hip::getCurrentHIPStream(image.device().index()).stream();
  const float *ip = image.data_ptr<float>();
  const float *mp = mask.data_ptr<float>();
  float *op = out.data_ptr<float>();
  void *wp = ws.data_ptr();
  int rc;
  if (c.nd == 2) {
    if (c.op == FG_OP_UNSIGNED)
      rc = fg_geodesic2d_f32(ip, mp, op, C, Cm, H, W, c.v, c.lg, c.le, c.iterations, wp, wsb, stream);
    else if (c.op == FG_OP_SIGNED)
      rc = fg_signed_geodesic2d_f32(ip, mp, op, C, Cm, H, W, c.v, c.lg, c.le, c.iterations, wp, wsb, stream);
    else
      rc = fg_gsf2d_f32(ip, mp, op, C, Cm, H, W, c.theta, c.v, c.lg, c.iterations, wp, wsb, stream);
  } else {
    const float *sp = c.spacing.data();
    if (c.op == FG_OP_UNSIGNED)
      rc = fg_geodesic3d_f32(ip, mp, op, C, Cm, D, H, W, sp, c.v, c.lg, c.le, c.iterations, wp, wsb, stream);
    else if (c.op == FG_OP_SIGNED)
      rc = fg_signed_geodesic3d_f32(ip, mp, op, C, Cm, D, H, W, sp, c.v, c.lg, c.le, c.iterations, wp, wsb,
                                    stream);
    else
      rc = fg_gsf3d_f32(ip, mp, op, C, Cm, D, H, W, c.theta, sp, c.v, c.lg, c.iterations, wp, wsb, stream);
  }
  raise_for(rc);
  if (sync_check_enabled()) raise_for(fg_status(stream));
  return out;
}

torch::Tensor dispatch(const Call &c, const torch::Tensor &image, const torch::Tensor &mask) {
  if (image.is_cuda()) {
    check_cuda(mask);
    return run_device(c, image, mask);
  }
  check_cpu(mask);
  // CPU tensors: compute on the current HIP device and hand the result back.
  if (!torch::cuda::is_available())
    throw std::runtime_error(
        "fastgeodis_amd computes on a ROCm GPU and none is available; the reference's OpenMP CPU path is "
        "not part of this library");
  const auto dev = torch::Device(torch::kCUDA, (c10::DeviceIndex)c10::hip::current_device());
  torch::Tensor out = run_device(c, image.to(dev), mask.to(dev));
  return out.to(torch::kCPU);
}

// ---- reference-compatible entry points (fastgeodis.cpp) -------------------

torch::Tensor generalised_geodesic2d(const torch::Tensor &image, const torch::Tensor &mask, const float &v,
                                     const float &l_grad, const float &l_eucl, const int &iterations) {
  check_input_dimensions(image, mask, 4);
  return dispatch({FG_OP_UNSIGNED, 2, {}, v, l_grad, l_eucl, 0.f, iterations}, image, mask);
}

torch::Tensor generalised_geodesic3d(const torch::Tensor &image, const torch::Tensor &mask,
                                     const std::vector<float> &spacing, const float &v, const float &l_grad,
                                     const float &l_eucl, const int &iterations) {
  check_input_dimensions(image, mask, 5);
  if (spacing.size() != 3)
    throw std::invalid_argument("function only supports 3D spacing inputs, received " +
                                std::to_string(spacing.size()));
  return dispatch({FG_OP_UNSIGNED, 3, spacing, v, l_grad, l_eucl, 0.f, iterations}, image, mask);
}

torch::Tensor signed_generalised_geodesic2d(const torch::Tensor &image, const torch::Tensor &mask,
                                            const float &v, const float &l_grad, const float &l_eucl,
                                            const int &iterations) {
  check_input_dimensions(image, mask, 4);
  return dispatch({FG_OP_SIGNED, 2, {}, v, l_grad, l_eucl, 0.f, iterations}, image, mask);
}

torch::Tensor signed_generalised_geodesic3d(const torch::Tensor &image, const torch::Tensor &mask,
                                            const std::vector<float> &spacing, const float &v,
                                            const float &l_grad, const float &l_eucl, const int &iterations) {
  check_input_dimensions(image, mask, 5);
  if (spacing.size() != 3)
    throw std::invalid_argument("function only supports 3D spacing inputs, received " +
                                std::to_string(spacing.size()));
  return dispatch({FG_OP_SIGNED, 3, spacing, v, l_grad, l_eucl, 0.f, iterations}, image, mask);
}

torch::Tensor GSF2d(const torch::Tensor &image, const torch::Tensor &mask, const float &theta, const float &v,
                    const float &lambda, const int &iterations) {
  check_input_dimensions(image, mask, 4);
  return dispatch({FG_OP_GSF, 2, {}, v, lambda, 0.f, theta, iterations}, image, mask);
}

torch::Tensor GSF3d(const torch::Tensor &image, const torch::Tensor &mask, const float &theta,
                    const std::vector<float> &spacing, const float &v, const float &lambda,
                    const int &iterations) {
  check_input_dimensions(image, mask, 5);
  if (spacing.size() != 3)
    throw std::invalid_argument("function only supports 3D spacing inputs, received " +
                                std::to_string(spacing.size()));
  return dispatch({FG_OP_GSF, 3, spacing, v, lambda, 0.f, theta, iterations}, image, mask);
}

// ---- extensions of this library -------------------------------------------

// Batched form: every instance is (image (1,C,[D,]H,W), mask (1,Cm,[D,]H,W)),
// all on one HIP device.  Returns one output per instance; each equals the
// single-instance call with the same arguments.
std::vector<torch::Tensor> geodesic_batch(const std::vector<torch::Tensor> &images,
                                          const std::vector<torch::Tensor> &masks,
                                          const std::vector<std::vector<float>> &spacings,
                                          const std::vector<int64_t> &ops, const std::vector<double> &vs,
                                          const std::vector<double> &lgs, const std::vector<double> &les,
                                          const std::vector<double> &thetas,
                                          const std::vector<int64_t> &iterations,
                                          const std::vector<torch::Tensor> &images_t) {
  const size_t n = images.size();
  if (masks.size() != n || spacings.size() != n || ops.size() != n || vs.size() != n || lgs.size() != n ||
      les.size() != n || thetas.size() != n || iterations.size() != n || (!images_t.empty() && images_t.size() != n))
    throw std::invalid_argument("batch argument lists must have equal lengths");
  std::vector<torch::Tensor> outs;
  if (n == 0) return outs;
  std::vector<torch::Tensor> ims, mks;
  std::vector<fg_instance> inst(n);
  const auto dev = images[0].device();
  if (!images[0].is_cuda()) throw std::invalid_argument("batched calls take HIP tensors");
  c10::DeviceGuard guard(dev);
  std::vector<fg_instance> work;  // instances with a non-empty output
  for (size_t i = 0; i < n; ++i) {
    const int nd = (int)images[i].dim() - 2;
    if (nd != 2 && nd != 3) throw std::invalid_argument("instance " + std::to_string(i) + ": need 2D or 3D");
    check_input_dimensions(images[i], masks[i], nd + 2);
    check_cuda(masks[i]);
    if (images[i].device() != dev || masks[i].device() != dev)
      throw std::invalid_argument("all batch tensors must be on one device");
    check_float(images[i], "image");
    check_float(masks[i], "mask");
    if (nd == 3 && spacings[i].size() != 3)
      throw std::invalid_argument("function only supports 3D spacing inputs, received " +
                                  std::to_string(spacings[i].size()));
    ims.push_back(images[i].contiguous());
    mks.push_back(masks[i].contiguous());
    outs.push_back(torch::empty_like(mks.back(), torch::MemoryFormat::Contiguous));
    if (outs.back().numel() == 0) continue;  // nothing to compute, as in the single-instance call
    if (ims[i].size(1) == 0) throw std::invalid_argument("instance " + std::to_string(i) + ": image must have at least one channel");
    fg_instance &f = inst[i];
    f.image = ims[i].data_ptr<float>();
    f.mask = mks[i].data_ptr<float>();
    f.out = outs[i].data_ptr<float>();
    f.C = (int)ims[i].size(1);
    f.Cm = (int)mks[i].size(1);
    f.is2d = nd == 2;
    f.D = nd == 2 ? 1 : (int)ims[i].size(2);
    f.H = (int)ims[i].size(nd == 2 ? 2 : 3);
    f.W = (int)ims[i].size(nd == 2 ? 3 : 4);
    f.op = (int)ops[i];
    f.iterations = (int)iterations[i];
    for (int k = 0; k < 3; ++k) f.spacing[k] = nd == 3 ? spacings[i][k] : 1.0f;
    f.v = (float)vs[i];
    f.l_grad = (float)lgs[i];
    f.l_eucl = (float)les[i];
    f.theta = (float)thetas[i];
    // an optional transposed image: (1, C, W, H, D) / (1, C, W, H), contiguous float32 on the same device
    if (!images_t.empty() && images_t[i].defined() && images_t[i].numel() > 0) {
      const torch::Tensor &t = images_t[i];
      check_float(t, "transposed image");
      std::vector<int64_t> want{1, ims[i].size(1)};
      for (int k = nd + 1; k >= 2; --k) want.push_back(ims[i].size(k));
      if (t.device() != dev || !t.is_contiguous() || t.sizes().vec() != want)
        throw std::invalid_argument("instance " + std::to_string(i) +
                                    ": transposed image must be a contiguous (1, C, W, H[, D]) tensor on the image's device");
      f.imageT = t.data_ptr<float>();
    }
    work.push_back(f);
  }
  if (work.empty()) return outs;
  const size_t wsb = fg_batch_workspace_bytes(work.data(), (int)work.size());
  if (wsb == 0) raise_for(FG_EINVAL);
  torch::Tensor ws = torch::empty({(int64_t)wsb}, ims[0].options().dtype(torch::kUInt8));
  void *stream = (void *)c10::hip::getCurrentHIPStream(dev.index()).stream();
  raise_for(fg_geodesic_batch_f32(work.data(), (int)work.size(), ws.data_ptr(), wsb, stream));
  if (sync_check_enabled()) raise_for(fg_status(stream));
  return outs;
}

// Writes values[i] at flat element indices[i] of masks[i] (float32 HIP tensors
// of one device) on that device's current stream.
void set_voxels(const std::vector<torch::Tensor> &masks, const std::vector<int64_t> &indices,
                const std::vector<double> &values) {
  const size_t n = masks.size();
  if (indices.size() != n || values.size() != n) throw std::invalid_argument("argument lists must have equal lengths");
  if (n == 0) return;
  const auto dev = masks[0].device();
  std::vector<fg_voxel_set> items(n);
  for (size_t i = 0; i < n; ++i) {
    const torch::Tensor &m = masks[i];
    check_cuda(m);
    check_float(m, "mask");
    if (m.device() != dev || !m.is_contiguous()) throw std::invalid_argument("masks must be contiguous and on one device");
    if (indices[i] < 0 || indices[i] >= m.numel()) throw std::invalid_argument("voxel index out of range");
    items[i].dst = m.data_ptr<float>();
    items[i].index = indices[i];
    items[i].value = (float)values[i];
  }
  c10::DeviceGuard guard(dev);
  void *stream = (void *)c10::hip::getCurrentHIPStream(dev.index()).stream();
  raise_for(fg_set_voxels_f32(items.data(), (int)n, stream));
}

void set_tuning(int engine, int halo, int max_groups, int shape) {
  fg_tuning t{engine, halo, max_groups, shape};
  raise_for(fg_tuning_set(&t));
}

std::vector<int64_t> get_tuning() {
  fg_tuning t{};
  raise_for(fg_tuning_get(&t));
  return {t.engine, t.halo, t.max_groups, t.shape};
}

// mode: False/0 off, True/1 launch events, 2 exchange-wait histogram, 3 both
void profile_enable(int64_t mode) { raise_for(fg_profile_enable((int)mode)); }

std::tuple<int64_t, double, double> profile_collect() {
  int n = 0;
  double ms = 0, bytes = 0;
  raise_for(fg_profile_collect(&n, &ms, &bytes));
  return {n, ms, bytes};
}

// Synchronises the current stream of `device` and raises if a launch on that
// device left an error in its status word since the last check.
void check_status(int64_t device) {
  c10::DeviceGuard guard(torch::Device(torch::kCUDA, (c10::DeviceIndex)device));
  void *stream = (void *)c10::hip::getCurrentHIPStream((c10::DeviceIndex)device).stream();
  raise_for(fg_status(stream));
}

int64_t pending_status(int64_t device) { return fg_pending_status((int)device); }

void inject_status(int64_t device, int64_t code) { raise_for(fg_inject_status((int)device, (int)code)); }

void set_wait_limit(int64_t spins) { raise_for(fg_set_wait_limit((int)spins)); }

void set_graphs(bool on) { raise_for(fg_set_graphs(on ? 1 : 0)); }

// (counts per log2 bucket of device-clock ticks, device clock rate in kHz)
std::tuple<std::vector<int64_t>, int64_t> profile_exchange_histogram() {
  unsigned h[32] = {};
  raise_for(fg_profile_exchange_histogram(h, 32));
  int dev = 0, khz = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  return {std::vector<int64_t>(h, h + 32), (int64_t)khz};
}

// Continuous batching at pass granularity (fg_stepper_*).  The stepper owns
// its exchange and descriptor buffers; every added instance keeps its
// tensors (image, mask, output, workspace, transposed image) alive until a
// step reports it done.  Steps run on the device's current stream.
class Stepper {
 public:
  Stepper(int64_t halo_bytes, int64_t max_transforms, int64_t device) : dev_(device) {
    c10::DeviceGuard guard(c10::Device(c10::kCUDA, (c10::DeviceIndex)device));
    const auto opt = torch::TensorOptions().dtype(torch::kUInt8).device(torch::kCUDA, (c10::DeviceIndex)device);
    const size_t bb = fg_stepper_blob_bytes((int)max_transforms);
    halo_ = torch::empty({halo_bytes}, opt);
    blob_ = torch::empty({(int64_t)bb}, opt);
    raise_for(fg_stepper_create(&s_, halo_.data_ptr(), (size_t)halo_bytes, blob_.data_ptr(), bb));
  }
  ~Stepper() {
    if (s_) fg_stepper_destroy(s_);
  }
  Stepper(const Stepper &) = delete;
  Stepper &operator=(const Stepper &) = delete;

  // (ticket, output); ticket -1: the exchange buffer cannot take the instance
  // now (retry after a step has retired others)
  std::tuple<int64_t, torch::Tensor> add(const torch::Tensor &image_in, const torch::Tensor &mask_in,
                                         const std::vector<float> &spacing, int64_t op, double v, double lg,
                                         double le, int64_t iterations,
                                         const c10::optional<torch::Tensor> &image_t_opt) {
    const int nd = (int)image_in.dim() - 2;
    if (nd != 2 && nd != 3) throw std::invalid_argument("need a 2D or 3D instance");
    check_input_dimensions(image_in, mask_in, nd + 2);
    check_cuda(image_in);
    check_cuda(mask_in);
    check_float(image_in, "image");
    check_float(mask_in, "mask");
    if (image_in.device().index() != dev_ || mask_in.device().index() != dev_)
      throw std::invalid_argument("instance tensors must be on the stepper's device");
    if (nd == 3 && spacing.size() != 3)
      throw std::invalid_argument("function only supports 3D spacing inputs, received " + std::to_string(spacing.size()));
    c10::DeviceGuard guard(image_in.device());
    Held h;
    h.image = image_in.contiguous();
    h.mask = mask_in.contiguous();
    h.out = torch::empty_like(h.mask, torch::MemoryFormat::Contiguous);
    fg_instance f{};
    f.image = h.image.data_ptr<float>();
    f.mask = h.mask.data_ptr<float>();
    f.out = h.out.data_ptr<float>();
    f.C = (int)h.image.size(1);
    f.Cm = (int)h.mask.size(1);
    f.is2d = nd == 2;
    f.D = nd == 2 ? 1 : (int)h.image.size(2);
    f.H = (int)h.image.size(nd == 2 ? 2 : 3);
    f.W = (int)h.image.size(nd == 2 ? 3 : 4);
    f.op = (int)op;
    f.iterations = (int)iterations;
    for (int k = 0; k < 3; ++k) f.spacing[k] = nd == 3 ? spacing[k] : 1.0f;
    f.v = (float)v;
    f.l_grad = (float)lg;
    f.l_eucl = (float)le;
    if (image_t_opt.has_value() && image_t_opt->defined() && image_t_opt->numel() > 0) {
      const torch::Tensor &image_t = *image_t_opt;
      check_float(image_t, "transposed image");
      std::vector<int64_t> want{1, h.image.size(1)};
      for (int k = nd + 1; k >= 2; --k) want.push_back(h.image.size(k));
      if (image_t.device() != h.image.device() || !image_t.is_contiguous() || image_t.sizes().vec() != want)
        throw std::invalid_argument("transposed image must be a contiguous (1, C, W, H[, D]) tensor on the image's device");
      h.image_t = image_t;
      f.imageT = image_t.data_ptr<float>();
    }
    size_t halo_share = 0;
    const size_t wsb = fg_stepper_op_bytes(&f, &halo_share);
    if (wsb == 0) raise_for(FG_EINVAL);
    h.ws = torch::empty({(int64_t)wsb}, h.image.options().dtype(torch::kUInt8));
    int ticket = 0;
    const int rc = fg_stepper_add(s_, &f, h.ws.data_ptr(), wsb, &ticket);
    if (rc == FG_EWORKSPACE && halo_share > 0) return {-1, torch::Tensor()};
    raise_for(rc);
    held_[ticket] = h;
    return {ticket, h.out};
  }

  // one step on the current stream; the tickets of the instances it completed
  std::vector<int64_t> step() {
    c10::DeviceGuard guard(c10::Device(c10::kCUDA, (c10::DeviceIndex)dev_));
    std::vector<int> done(held_.size() + 1);
    int n = 0;
    void *stream = (void *)c10::hip::getCurrentHIPStream((c10::DeviceIndex)dev_).stream();
    raise_for(fg_stepper_step(s_, stream, done.data(), (int)done.size(), &n));
    std::vector<int64_t> out;
    for (int i = 0; i < n; ++i) {
      out.push_back(done[i]);
      held_.erase(done[i]);  // later work on this stream is ordered after the step
    }
    return out;
  }

  int64_t active() const { return fg_stepper_active(s_); }

 private:
  struct Held {
    torch::Tensor image, mask, out, ws, image_t;
  };
  int64_t dev_;
  torch::Tensor halo_, blob_;
  fg_stepper *s_ = nullptr;
  std::map<int, Held> held_;
};

int64_t workspace_bytes(int64_t op, int64_t C, int64_t D, int64_t H, int64_t W) {
  return (int64_t)fg_workspace_bytes((int)op, (int)C, (int)D, (int)H, (int)W);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.def("generalised_geodesic2d", &generalised_geodesic2d, "Generalised Geodesic distance 2d");
  m.def("generalised_geodesic3d", &generalised_geodesic3d, "Generalised Geodesic distance 3d");
  m.def("signed_generalised_geodesic2d", &signed_generalised_geodesic2d, "Signed Generalised Geodesic distance 2d");
  m.def("signed_generalised_geodesic3d", &signed_generalised_geodesic3d, "Signed Generalised Geodesic distance 3d");
  m.def("GSF2d", &GSF2d, "Geodesic Symmetric Filtering 2d");
  m.def("GSF3d", &GSF3d, "Geodesic Symmetric Filtering 3d");
  m.def("geodesic_batch", &geodesic_batch, "Batched raster-scan transforms of uneven sizes",
        pybind11::arg("images"), pybind11::arg("masks"), pybind11::arg("spacings"), pybind11::arg("ops"),
        pybind11::arg("vs"), pybind11::arg("lgs"), pybind11::arg("les"), pybind11::arg("thetas"),
        pybind11::arg("iterations"), pybind11::arg("images_t") = std::vector<torch::Tensor>());
  m.def("set_voxels", &set_voxels, "Write single mask voxels (one launch per 96)");
  m.def("set_tuning", &set_tuning, "Set engine / halo depth / workgroup cap / 3D shape for this host thread");
  m.def("get_tuning", &get_tuning, "Current (engine, halo, max_groups, shape) of this host thread");
  m.def("workspace_bytes", &workspace_bytes, "Workspace bytes of one call");
  m.def("abi_version", []() { return fg_abi_version(); });
  m.def("profile_enable", &profile_enable, "Bracket tiled sweep launches with HIP events");
  m.def("profile_collect", &profile_collect, "(launches, total_ms, algorithmic_bytes) since enabling");
  m.def("profile_exchange_histogram", &profile_exchange_histogram,
        "(counts per log2 bucket of device-clock ticks, clock kHz) of halo-exchange waits since enabling");
  m.def("check_status", &check_status, "Synchronise the device's current stream and raise on a pending device error");
  m.def("pending_status", &pending_status, "Read and clear a device's pending status code without synchronising");
  m.def("inject_status", &inject_status, "Set a device's pending status from the host (error-path tests)");
  m.def("set_wait_limit", &set_wait_limit, "Spin limit of the bounded halo waits for this host thread (0 = default)");
  m.def("release", []() { fg_release(); }, "Free the library's pinned memory, events and graphs (process exit)");
  m.def("set_graphs", &set_graphs, "Graph replay of repeated launch sequences on this host thread (default on)");
  pybind11::class_<Stepper>(m, "Stepper", "Continuous batching at pass granularity (unsigned and signed transforms)")
      .def(pybind11::init<int64_t, int64_t, int64_t>(), pybind11::arg("halo_bytes"), pybind11::arg("max_transforms"),
           pybind11::arg("device"))
      .def("add", &Stepper::add, "Add an instance: (ticket, output); ticket -1 when the exchange buffer is full",
           pybind11::arg("image"), pybind11::arg("mask"), pybind11::arg("spacing"), pybind11::arg("op"),
           pybind11::arg("v"), pybind11::arg("lg"), pybind11::arg("le"), pybind11::arg("iterations"),
           pybind11::arg("image_t") = pybind11::none())
      .def("step", &Stepper::step, "Enqueue one step on the current stream; tickets it completed")
      .def("active", &Stepper::active, "Instances in flight");
}
