// Python module `gpuNUFFT` with the reference's NUFFTOp
// (CUDA/src/gpu/python/gpuNUFFT_operator_python_factory.cpp:60-182), built on
// the C++ API of include/gpuNUFFT_operator_factory.hpp.
//
// Same constructor arguments, methods and memory conventions: array buffers
// are read as raw memory in the reference layouts (trajectory SoA [dim][M],
// k-space [coil][M], images [coil][z][y][x]); inputs of another dtype are
// converted by pybind11 as in the reference. Differences, all supersets:
// keyword arguments and defaults, a float `osr`, `sense_maps=None`, 2-D
// adjoint outputs of shape (coils, 1, H, W) instead of a zero-depth array,
// and grid outputs (grid_data=True) always per coil.

#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <complex>
#include <cstring>
#include <vector>

#include "gpuNUFFT_operator_factory.hpp"

namespace py = pybind11;

namespace {

template <typename T>
T* raw(py::array_t<T>& a) {
  return static_cast<T*>(a.request().ptr);
}

class PyNUFFTOp {
 public:
  PyNUFFTOp(py::array_t<float> kspace_loc, py::array_t<int> image_size, int num_coils,
            py::object sense_maps, py::array_t<float> density_comp, int kernel_width,
            int sector_width, float osr, bool balance_workload)
      : n_coils_(num_coils) {
    py::buffer_info loc = kspace_loc.request();
    if (loc.ndim != 2) throw std::invalid_argument("kspace_loc must have shape (dim, M)");
    dim_ = (int)loc.shape[0];
    M_ = (int)loc.shape[1];
    if (dim_ != 2 && dim_ != 3) throw std::invalid_argument("kspace_loc: dim must be 2 or 3");
    gpuNUFFT::Array<DType> traj;
    traj.data = static_cast<DType*>(loc.ptr);
    traj.dim.length = (IndType)M_;

    gpuNUFFT::Array<DType> dens;
    if (density_comp.size() > 0) {
      if (density_comp.size() < M_) throw std::invalid_argument("density_comp: fewer than M weights");
      dens.data = raw(density_comp);
      dens.dim.length = (IndType)M_;
    }

    py::buffer_info isz = image_size.request();
    if (isz.size < dim_) throw std::invalid_argument("image_size: fewer entries than dimensions");
    const int* d = static_cast<const int*>(isz.ptr);
    img_.width = (IndType)d[0];
    img_.height = (IndType)d[1];
    img_.depth = dim_ == 3 ? (IndType)d[2] : 0;

    gpuNUFFT::Array<DType2> sensArr;
    if (!sense_maps.is_none()) {
      auto s = py::array_t<std::complex<float>>::ensure(sense_maps);
      if (!s) throw std::invalid_argument("sense_maps: not convertible to complex64");
      if (s.ndim() > 0) {
        const size_t n = (size_t)img_.count() * (size_t)n_coils_;
        if ((size_t)s.size() < n) throw std::invalid_argument("sense_maps: fewer than coils*pixels");
        // copied, as the reference does: the operator reads them per call
        sens_.resize(n);
        std::memcpy(sens_.data(), s.request().ptr, n * sizeof(DType2));
        sensArr.data = sens_.data();
        sensArr.dim = img_;
        sensArr.dim.channels = (IndType)n_coils_;
      }
    }
    factory_.setBalanceWorkload(balance_workload);
    op_ = factory_.createGpuNUFFTOperator(traj, dens, sensArr, (IndType)kernel_width,
                                          (IndType)sector_width, (DType)osr, img_);
  }

  ~PyNUFFTOp() { delete op_; }

  py::array_t<std::complex<float>> op(py::array_t<std::complex<float>> image, bool interpolate) {
    py::array_t<std::complex<float>> out({(py::ssize_t)n_coils_, (py::ssize_t)M_});
    gpuNUFFT::Array<CufftType> k;
    k.data = reinterpret_cast<CufftType*>(out.request().ptr);
    k.dim.length = (IndType)M_;
    k.dim.channels = (IndType)n_coils_;
    const size_t need = (size_t)img_.count() * (sens_.empty() ? (size_t)n_coils_ : 1);
    if ((size_t)image.size() < need) throw std::invalid_argument("op: image has too few elements");
    gpuNUFFT::Array<DType2> im;
    im.data = reinterpret_cast<DType2*>(image.request().ptr);
    im.dim = img_;
    im.dim.channels = (IndType)n_coils_;
    {
      py::gil_scoped_release nogil;
      if (interpolate)
        op_->performForwardGpuNUFFT(im, k, gpuNUFFT::DENSITY_ESTIMATION);
      else
        op_->performForwardGpuNUFFT(im, k);
    }
    return out;
  }

  py::array_t<std::complex<float>> adj_op(py::array_t<std::complex<float>> kspace_data,
                                          bool grid_data) {
    gpuNUFFT::Dimensions dims = img_;
    if (grid_data) dims = dims * op_->getOsf();
    const py::ssize_t depth = dim_ == 3 ? (py::ssize_t)dims.depth : 1;
    const bool combined = !sens_.empty() && !grid_data;
    std::vector<py::ssize_t> shape;
    if (!combined) shape.push_back(n_coils_);
    shape.push_back(depth);
    shape.push_back((py::ssize_t)dims.height);
    shape.push_back((py::ssize_t)dims.width);
    py::array_t<std::complex<float>> out(shape);
    if ((size_t)kspace_data.size() < (size_t)M_ * n_coils_)
      throw std::invalid_argument("adj_op: kspace_data has too few elements");
    gpuNUFFT::Array<CufftType> img;
    img.data = reinterpret_cast<CufftType*>(out.request().ptr);
    img.dim = dims;
    if (!combined) img.dim.channels = (IndType)n_coils_;
    gpuNUFFT::Array<DType2> k;
    k.data = reinterpret_cast<DType2*>(kspace_data.request().ptr);
    k.dim.length = (IndType)M_;
    k.dim.channels = (IndType)n_coils_;
    {
      py::gil_scoped_release nogil;
      if (grid_data)
        op_->performGpuNUFFTAdj(k, img, gpuNUFFT::DENSITY_ESTIMATION);
      else
        op_->performGpuNUFFTAdj(k, img);
    }
    return out;
  }

  gpuNUFFT::GpuNUFFTOperator* cpp() { return op_; }

 private:
  gpuNUFFT::GpuNUFFTOperatorFactory factory_;
  gpuNUFFT::GpuNUFFTOperator* op_ = nullptr;
  int n_coils_ = 1, dim_ = 2, M_ = 0;
  gpuNUFFT::Dimensions img_;
  std::vector<DType2> sens_;
};

// Copy of the first `n` elements of a table (default: its count()).
template <typename T>
py::array_t<T> to_numpy(gpuNUFFT::Array<T> a, size_t n = (size_t)-1) {
  if (n == (size_t)-1) n = a.count();
  py::array_t<T> out((py::ssize_t)n);
  if (n) std::memcpy(out.request().ptr, a.data, n * sizeof(T));
  return out;
}

}  // namespace

PYBIND11_MODULE(gpuNUFFT, m) {
  m.doc() = "gpuNUFFT NUFFTOp on AMD Instinct (HIP), drop-in for the reference module";
  py::class_<PyNUFFTOp>(m, "NUFFTOp")
      .def(py::init<py::array_t<float>, py::array_t<int>, int, py::object, py::array_t<float>, int,
                    int, float, bool>(),
           py::arg("kspace_loc"), py::arg("image_size"), py::arg("num_coils"),
           py::arg("sense_maps") = py::none(), py::arg("density_comp") = py::array_t<float>(),
           py::arg("kernel_width") = 3, py::arg("sector_width") = 8, py::arg("osr") = 2.0f,
           py::arg("balance_workload") = true)
      .def("op", &PyNUFFTOp::op, py::arg("image"), py::arg("interpolate_data") = false)
      .def("adj_op", &PyNUFFTOp::adj_op, py::arg("kspace_data"), py::arg("grid_data") = false)
      // precomputed tables of the underlying C++ operator (getters of
      // CUDA/inc/gpuNUFFT_operator.hpp:136-201), for inspection and tests
      .def("get_tables",
           [](PyNUFFTOp& self) {
             gpuNUFFT::GpuNUFFTOperator* o = self.cpp();
             py::dict d;
             d["data_indices"] = to_numpy(o->getDataIndices());
             d["sector_data_count"] = to_numpy(o->getSectorDataCount());
             d["sector_centers"] = to_numpy(o->getSectorCenters());
             // SoA [dim][M]; dim.length holds M (reference convention)
             gpuNUFFT::Array<DType> traj = o->getKSpaceTraj();
             d["kspace_traj"] = to_numpy(traj, (size_t)o->getImageDimensionCount() * traj.count());
             d["kernel"] = to_numpy(o->getKernel());
             d["deapo"] = to_numpy(o->getDeapodizationFunction());
             d["operator_type"] = (int)o->getType();
             if (auto* b = dynamic_cast<gpuNUFFT::BalancedOperator*>(o)) {
               gpuNUFFT::Array<IndType2> order = b->getSectorProcessingOrder();
               py::array_t<unsigned int> a({(py::ssize_t)order.count(), (py::ssize_t)2});
               if (order.count())
                 std::memcpy(a.request().ptr, order.data, order.count() * sizeof(IndType2));
               d["sector_processing_order"] = a;
             }
             return d;
           });
}
