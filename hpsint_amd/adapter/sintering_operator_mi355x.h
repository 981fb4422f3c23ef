// sintering_operator_mi355x.h -- the C++ adapter through which hpsint's
// Newton-Krylov loop drives the MI355X operator (header-only, C++17, host
// code over the C ABI of include/hpsint_amd.h).
//
// It implements the nonlinear-operator concept that TimeMarchingImplicit
// (include/pf-applications/time_integration/time_marching.h:129-241) and
// JacobianWrapper / ResidualWrapperGeneric (include/pf-applications/lac/
// jacobian_operator.h:66-76, residual_wrapper.h:179-223) call on
// SinteringOperatorGeneric (applications/sintering/include/pf-applications/
// sintering/operator_sintering_generic.h:341-677):
//   evaluate_nonlinear_residual<wtd>(dst, src, pre, post)   (:393-401)
//   vmult(dst, src)                                          (operator_base.h:294)
//   initialize_dof_vector, n_components, n_grains, equation_type
// plus the data calls the driver makes between them
// (SinteringOperatorData::fill_quadrature_point_values, set_component_mask;
// TimeIntegratorData weights) and the preconditioner concept
// (PreconditionerBase::do_update / vmult, include/pf-applications/lac/
// preconditioners.h:77-112).
//
// Template arguments are hpsint's own types:
//   Vector   a device-resident block vector; data() is the interleaved
//            device buffer of n_nodes * n_components doubles (the layout of
//            VectorTools::merge_components_fast)
//   History  provides get_old_solutions() -> a range of const Vector *
//            (SolutionHistory, time_integration/solution_history.h:186-190)
// Status codes become exceptions: HPS_ERR_NEWTON_NO_CONVERGENCE ->
// NewtonDidNotConverge, HPS_ERR_INVALID_N_COMP -> InvalidNumberOfComponents,
// anything else -> std::runtime_error; hpsint maps the first two to its
// ExcNewtonDidNotConverge / ExcInvalidNumberOfComponents where it catches
// them (driver.h:2489-2507).
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hpsint_amd.h"

namespace hpsint_mi355x
{
  struct NewtonDidNotConverge : std::runtime_error
  {
    using std::runtime_error::runtime_error;
  };
  struct InvalidNumberOfComponents : std::runtime_error
  {
    using std::runtime_error::runtime_error;
  };

  inline void check(int rc)
  {
    if (rc == HPS_OK)
      return;
    const std::string msg = hps_last_error();
    if (rc == HPS_ERR_NEWTON_NO_CONVERGENCE)
      throw NewtonDidNotConverge(msg);
    if (rc == HPS_ERR_INVALID_N_COMP)
      throw InvalidNumberOfComponents(msg);
    throw std::runtime_error(msg);
  }

  // EquationType of operator_sintering_generic.h:634-639: c and the order
  // parameters carry a time derivative, mu is stationary
  enum class EquationType
  {
    TimeDependent,
    Stationary
  };

  template <typename Vector, typename History>
  class SinteringOperatorMI355X
  {
  public:
    using value_type = double;
    // the range hook of evaluate_nonlinear_residual: called with
    // [begin, end) of the locally owned entries before (pre) and after
    // (post) the cells that touch them; the GPU evaluates all cells in one
    // launch, so the range is the whole vector, pre before the launch and
    // post after it (residual_traits.h:22-38 allows any partition)
    using RangeHook = std::function<void(unsigned int, unsigned int)>;

    SinteringOperatorMI355X(const hps_params &params, const History &history)
      : history(history)
    {
      check(hps_create(&params, &op));
      check(hps_sizes(op, &n_nodes_, &n_dofs_));
    }
    ~SinteringOperatorMI355X() { hps_destroy(op); }
    SinteringOperatorMI355X(const SinteringOperatorMI355X &) = delete;
    SinteringOperatorMI355X &operator=(const SinteringOperatorMI355X &) = delete;

    // OperatorBase::n_components / n_grains / n_grains_to_n_components
    unsigned int n_components() const { return (unsigned int)(n_dofs_ / n_nodes_); }
    unsigned int n_grains() const { return n_components() - 2; }
    static unsigned int n_grains_to_n_components(unsigned int n_grains)
    {
      return n_grains + 2;
    }
    EquationType equation_type(unsigned int component) const
    {
      return component == 1 ? EquationType::Stationary : EquationType::TimeDependent;
    }
    // OperatorBase::initialize_dof_vector: the size hpsint allocates on the
    // device for one block vector
    int64_t n_dofs() const { return n_dofs_; }
    int64_t n_nodes() const { return n_nodes_; }

    // TimeIntegratorData: primary weight, algebraic weight and the weights
    // of the old solutions (time_integrator_data.h:152-170)
    void set_time_weights(double w0, double w_alg, const std::vector<double> &w_hist)
    {
      check(hps_set_time_weights(op, w0, w_alg, w_hist.data(), (int)w_hist.size()));
      order = (int)w_hist.size();
    }

    // SinteringOperatorData::fill_quadrature_point_values
    void fill_quadrature_point_values(const Vector &u) { check(hps_set_linearization(op, u.data())); }

    // SinteringOperatorData::set_component_mask with the driver's grain
    // cut-off tolerance (driver.h:2212-2218); tol <= 0 switches it off
    void set_grain_cut_off(double tol) { check(hps_set_grain_cut_off(op, tol)); }
    void set_component_mask(const Vector &u) { check(hps_set_component_mask(op, u.data())); }

    // constrained DoFs of MatrixFree::get_constrained_dofs, as node indices
    void set_constrained_nodes(const std::vector<int64_t> &nodes)
    {
      check(hps_set_constrained_nodes(op, nodes.data(), (int64_t)nodes.size()));
    }

    template <int with_time_derivative = 2>
    void evaluate_nonlinear_residual(Vector &dst, const Vector &src, const RangeHook &pre = {},
                                     const RangeHook &post = {}) const
    {
      std::vector<const double *> old;
      if (with_time_derivative == 2)
        for (const Vector *v : history.get_old_solutions())
          if ((int)old.size() < order)
            old.push_back(v->data());
      if (pre)
        pre(0, (unsigned int)n_dofs_);
      check(hps_residual(op, with_time_derivative, dst.data(), src.data(),
                         old.empty() ? nullptr : old.data()));
      if (post)
        {
          // the hook runs on the host: the residual must be complete
          check(hps_sync(op));
          post(0, (unsigned int)n_dofs_);
        }
    }

    void vmult(Vector &dst, const Vector &src) const { check(hps_vmult_f64(op, dst.data(), src.data())); }

    // PreconditionerBase on the same operator: kind = HPS_PRECON_*
    void precondition_update(int kind) { check(hps_precond_update(op, kind)); }
    void precondition_vmult(Vector &dst, const Vector &src) const
    {
      check(hps_precond_apply_f64(op, dst.data(), src.data()));
    }

    void sync() const { check(hps_sync(op)); }
    hps_operator handle() const { return op; }

  private:
    const History &history;
    hps_operator   op      = nullptr;
    int64_t        n_nodes_ = 0, n_dofs_ = 0;
    int            order   = 0;
  };
} // namespace hpsint_mi355x
