// adapter_check.cpp -- a C entry point per adapter call, so that the adapter
// of sintering_operator_mi355x.h is compiled against include/hpsint_amd.h and
// exercised in-process (tests/test_gpu_adapter.py) with device buffers from
// the caller.  Every call waits for its work on the operator's stream.  The
// vector and history types are minimal stand-ins for hpsint's device block
// vector and SolutionHistory.
#include <cstdint>
#include <exception>
#include <string>
#include <vector>

#include "sintering_operator_mi355x.h"

namespace
{
  struct DeviceVector
  {
    double       *ptr;
    double       *data() { return ptr; }
    const double *data() const { return ptr; }
  };

  struct History
  {
    std::vector<const DeviceVector *> old;
    const std::vector<const DeviceVector *> &get_old_solutions() const { return old; }
  };

  struct Check
  {
    History                                                         history;
    std::vector<DeviceVector>                                       old_vectors;
    hpsint_mi355x::SinteringOperatorMI355X<DeviceVector, History> op;
    unsigned int pre_calls = 0, post_calls = 0, last_end = 0;
    Check(const hps_params &p)
      : op(p, history)
    {}
  };

  std::string g_error;

  template <typename F>
  int guarded(F &&f)
  {
    try
      {
        f();
        return 0;
      }
    catch (const hpsint_mi355x::NewtonDidNotConverge &e)
      {
        g_error = e.what();
        return HPS_ERR_NEWTON_NO_CONVERGENCE;
      }
    catch (const hpsint_mi355x::InvalidNumberOfComponents &e)
      {
        g_error = e.what();
        return HPS_ERR_INVALID_N_COMP;
      }
    catch (const std::exception &e)
      {
        g_error = e.what();
        return -1;
      }
  }
} // namespace

extern "C"
{
  const char *hpsad_last_error() { return g_error.c_str(); }

  int hpsad_create(const hps_params *p, void **out)
  {
    return guarded([&] { *out = new Check(*p); });
  }

  int hpsad_destroy(void *h)
  {
    delete static_cast<Check *>(h);
    return 0;
  }

  int hpsad_n_components(void *h) { return (int)static_cast<Check *>(h)->op.n_components(); }

  int hpsad_set_time(void *h, double w0, double w_alg, const double *w_hist, int order,
                     double *const *old)
  {
    return guarded([&] {
      Check *c = static_cast<Check *>(h);
      c->op.set_time_weights(w0, w_alg, std::vector<double>(w_hist, w_hist + order));
      c->old_vectors.clear();
      for (int i = 0; i < order; ++i)
        c->old_vectors.push_back(DeviceVector{old[i]});
      c->history.old.clear();
      for (const auto &v : c->old_vectors)
        c->history.old.push_back(&v);
    });
  }

  int hpsad_fill(void *h, double *u)
  {
    return guarded([&] {
      Check *c = static_cast<Check *>(h);
      c->op.fill_quadrature_point_values(DeviceVector{u});
      c->op.sync();
    });
  }

  int hpsad_vmult(void *h, double *dst, double *src)
  {
    return guarded([&] {
      Check       *c = static_cast<Check *>(h);
      DeviceVector d{dst};
      c->op.vmult(d, DeviceVector{src});
      c->op.sync();
    });
  }

  // residual with counting range hooks
  int hpsad_residual(void *h, int wtd, double *dst, double *src)
  {
    return guarded([&] {
      Check       *c = static_cast<Check *>(h);
      DeviceVector d{dst};
      auto pre  = [c](unsigned int, unsigned int) { ++c->pre_calls; };
      auto post = [c](unsigned int, unsigned int e) {
        ++c->post_calls;
        c->last_end = e;
      };
      if (wtd == 0)
        c->op.evaluate_nonlinear_residual<0>(d, DeviceVector{src}, pre, post);
      else if (wtd == 1)
        c->op.evaluate_nonlinear_residual<1>(d, DeviceVector{src}, pre, post);
      else
        c->op.evaluate_nonlinear_residual<2>(d, DeviceVector{src}, pre, post);
      c->op.sync();
    });
  }

  int hpsad_hook_calls(void *h, unsigned int *pre, unsigned int *post, unsigned int *end)
  {
    Check *c = static_cast<Check *>(h);
    *pre     = c->pre_calls;
    *post    = c->post_calls;
    *end     = c->last_end;
    return 0;
  }

  int hpsad_precond(void *h, int kind, double *dst, double *src)
  {
    return guarded([&] {
      Check *c = static_cast<Check *>(h);
      c->op.precondition_update(kind);
      DeviceVector d{dst};
      c->op.precondition_vmult(d, DeviceVector{src});
      c->op.sync();
    });
  }
}
