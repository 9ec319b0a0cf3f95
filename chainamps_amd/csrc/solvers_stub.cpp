#include "solvers.hpp"
namespace cam {
Ctx::~Ctx() {
  if (stream) (void)hipStreamDestroy(stream);
}
void heff_lowest(Heff&, const cplx*, bool, double, int, bool, double*, cplx*, double*) { CAM_THROW("not implemented"); }
void svd_truncate(Ctx&, const BSTensor&, int64_t, double, TensorPtr&, TensorPtr&, std::vector<double>&, double&) { CAM_THROW("not implemented"); }
void tebd_gate(Ctx&, const BSTensor&, const BSTensor&, const BSTensor&, const std::vector<double>&, int64_t, double, TensorPtr&, TensorPtr&, std::vector<double>&, double&) { CAM_THROW("not implemented"); }
void heff_vjp(Heff&, const cplx*, cplx*, cplx*, int64_t) { CAM_THROW("not implemented"); }
}
