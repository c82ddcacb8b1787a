// E_RHF(R) as a libtorch autograd Function, registered as the operator
// torch.ops.m2.rhf_energy.
//
// forward:  SCF::rhf (src/scf.cpp:90-185) on the device through m2_rhf at the
//           nuclear positions R (natom x 3, bohr); returns E (0-dim FP64) and
//           keeps D = 2 C_occ C_occ^T and W = C_occ diag(eps_occ) C_occ^T.
//           guess 1 starts from the superposition of atomic densities over
//           the minimal (STO-3G) shells given as (lm, nprim_m, exps_m,
//           coefs_m, atoms_m), the reference's default
//           (src/ProgramController.cpp:307); guess 0 from the core
//           Hamiltonian (the minimal tensors may then be empty).
// backward: dE/dR = Fock::compute_forces (src/rhfgrad.cpp:15-89) from D and
//           W through m2_grad, whose two-electron part is the derivative-ERI
//           adjoint kernel (integrals recomputed, never stored), scaled by
//           the incoming gradient.  Shell centres follow their atoms, so the
//           basis is rebuilt at R before the gradient call.
//
// The function is first-order only (backward is not differentiable).  Each
// device holds one engine context for this operator; calls on one device are
// serialised.
#include <torch/library.h>
#include <torch/torch.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "m2_scf.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

std::mutex g_mu;
std::map<int, m2_ctx*> g_ctx;

m2_ctx* context_for(int device) {
    auto it = g_ctx.find(device);
    if (it != g_ctx.end()) return it->second;
    m2_ctx* c = nullptr;
    const m2_status st = m2_ctx_create(&c, device);
    TORCH_CHECK(st == 0 && c, "m2.rhf_energy: no engine context on device ", device, ": ",
                c ? m2_last_error(c) : "m2_ctx_create failed");
    g_ctx[device] = c;
    return c;
}

// host copies of the inputs of one molecule
struct Host {
    int natom = 0, nsh = 0, n = 0;
    std::vector<double> Z, xyz, exps, coefs, centers;
    std::vector<int> l, nprim, atoms;
};

Host host_inputs(const Tensor& R, const Tensor& Z, const Tensor& l, const Tensor& nprim, const Tensor& exps,
                 const Tensor& coefs, const Tensor& atoms, bool pure) {
    TORCH_CHECK(R.dim() == 2 && R.size(1) == 3, "m2.rhf_energy: R must be natom x 3");
    Host h;
    h.natom = (int)R.size(0);
    TORCH_CHECK(Z.numel() == h.natom, "m2.rhf_energy: Z must hold natom charges");
    h.nsh = (int)l.numel();
    TORCH_CHECK(nprim.numel() == h.nsh && atoms.numel() == h.nsh, "m2.rhf_energy: l, nprim, atoms differ in length");
    auto cpu_d = [](const Tensor& t) { return t.detach().to(torch::kCPU, torch::kFloat64).contiguous(); };
    auto cpu_i = [](const Tensor& t) { return t.detach().to(torch::kCPU, torch::kInt32).contiguous(); };
    const Tensor r = cpu_d(R), z = cpu_d(Z), e = cpu_d(exps), c = cpu_d(coefs);
    const Tensor li = cpu_i(l), pi = cpu_i(nprim), ai = cpu_i(atoms);
    h.xyz.assign(r.data_ptr<double>(), r.data_ptr<double>() + r.numel());
    h.Z.assign(z.data_ptr<double>(), z.data_ptr<double>() + z.numel());
    h.l.assign(li.data_ptr<int>(), li.data_ptr<int>() + h.nsh);
    h.nprim.assign(pi.data_ptr<int>(), pi.data_ptr<int>() + h.nsh);
    h.atoms.assign(ai.data_ptr<int>(), ai.data_ptr<int>() + h.nsh);
    long nprim_tot = 0;
    for (int s = 0; s < h.nsh; ++s) {
        TORCH_CHECK(h.nprim[s] > 0 && h.l[s] >= 0, "m2.rhf_energy: bad shell ", s);
        TORCH_CHECK(h.atoms[s] >= 0 && h.atoms[s] < h.natom, "m2.rhf_energy: shell ", s, " names atom ", h.atoms[s],
                    " of ", h.natom);
        nprim_tot += h.nprim[s];
        h.n += pure ? 2 * h.l[s] + 1 : (h.l[s] + 1) * (h.l[s] + 2) / 2;
    }
    TORCH_CHECK(e.numel() == nprim_tot && c.numel() == nprim_tot,
                "m2.rhf_energy: exps/coefs must hold sum(nprim) values");
    h.exps.assign(e.data_ptr<double>(), e.data_ptr<double>() + nprim_tot);
    h.coefs.assign(c.data_ptr<double>(), c.data_ptr<double>() + nprim_tot);
    h.centers.resize(3 * h.nsh);
    for (int s = 0; s < h.nsh; ++s)
        for (int k = 0; k < 3; ++k) h.centers[3 * s + k] = h.xyz[3 * h.atoms[s] + k];
    return h;
}

class RHFEnergy : public torch::autograd::Function<RHFEnergy> {
   public:
    static Tensor forward(AutogradContext* ctx, const Tensor& R, const Tensor& Z, const Tensor& l,
                          const Tensor& nprim, const Tensor& exps, const Tensor& coefs, const Tensor& atoms,
                          int64_t pure, int64_t charge, int64_t multiplicity, double precision, double enconverge,
                          double densconverge, int64_t maxiter, int64_t mixed, int64_t guess, const Tensor& lm,
                          const Tensor& nprim_m, const Tensor& exps_m, const Tensor& coefs_m, const Tensor& atoms_m) {
        Host h = host_inputs(R, Z, l, nprim, exps, coefs, atoms, pure != 0);
        TORCH_CHECK(guess == 0 || guess == 1, "m2.rhf_energy: guess must be 0 (core) or 1 (SOAD)");
        Host hm;
        if (guess == 1) {
            TORCH_CHECK(lm.numel() > 0, "m2.rhf_energy: the SOAD guess needs the minimal (STO-3G) shells");
            hm = host_inputs(R, Z, lm, nprim_m, exps_m, coefs_m, atoms_m, pure != 0);
        }
        const int device = R.is_cuda() ? R.get_device() : 0;
        const int n = h.n;
        std::vector<double> D((size_t)n * n), C((size_t)n * n), eps(n);
        m2_scf_options_c o{};
        o.diis = 1;
        o.maxdiis = 8;
        o.maxiter = (int)maxiter;
        o.enconverge = enconverge;
        o.densconverge = densconverge;
        o.precision = precision;
        o.df = 0;
        o.mixed = (int)mixed;
        o.mixed_threshold = 1e-4;
        o.fp64_final = 2;
        o.guess = (int)guess;
        o.diis_start = 2;
        m2_scf_result_c res{};
        {
            std::lock_guard<std::mutex> lk(g_mu);
            m2_ctx* c = context_for(device);
            const int st = m2_rhf(c, h.natom, h.Z.data(), h.xyz.data(), (int)charge, (int)multiplicity, (int)pure,
                                  h.nsh, h.l.data(), h.nprim.data(), h.exps.data(), h.coefs.data(), h.centers.data(),
                                  h.atoms.data(), 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, hm.nsh,
                                  hm.l.data(), hm.nprim.data(), hm.exps.data(), hm.coefs.data(), hm.centers.data(),
                                  hm.atoms.data(), &o, &res, D.data(), C.data(), eps.data(), nullptr, nullptr);
            TORCH_CHECK(st == 0, "m2.rhf_energy: ", m2_rhf_last_error());
        }
        TORCH_CHECK(res.converged, "m2.rhf_energy: SCF did not converge in ", maxiter, " iterations");
        // W = C_occ diag(eps_occ) C_occ^T (src/rhfgrad.cpp:35-46); C is column-major
        int nel = -(int)charge;
        for (double z : h.Z) nel += (int)z;
        const int nocc = nel / 2;
        auto opt = torch::TensorOptions().dtype(torch::kFloat64);
        Tensor Dt = torch::from_blob(D.data(), {n, n}, opt).clone();
        Tensor Ct = torch::from_blob(C.data(), {n, n}, opt).t().contiguous();   // row-major view of C
        Tensor Co = Ct.narrow(1, 0, nocc);
        Tensor Wt = torch::matmul(Co * torch::from_blob(eps.data(), {n}, opt).narrow(0, 0, nocc), Co.t());
        ctx->save_for_backward({R.detach(), Z.detach(), l, nprim, exps, coefs, atoms, Dt, Wt.contiguous()});
        ctx->saved_data["pure"] = pure;
        ctx->saved_data["precision"] = precision;
        return torch::full({}, res.energy, R.options().dtype(torch::kFloat64));
    }

    static variable_list backward(AutogradContext* ctx, variable_list grad_out) {
        const auto s = ctx->get_saved_variables();
        const Tensor& R = s[0];
        const bool pure = ctx->saved_data["pure"].toInt() != 0;
        const double precision = ctx->saved_data["precision"].toDouble();
        Host h = host_inputs(R, s[1], s[2], s[3], s[4], s[5], s[6], pure);
        const Tensor D = s[7].contiguous(), W = s[8].contiguous();
        TORCH_CHECK(D.size(0) == h.n && W.size(0) == h.n, "m2.rhf_energy: saved densities do not fit the basis");
        const int device = R.is_cuda() ? R.get_device() : 0;
        std::vector<double> g(3 * (size_t)h.natom);
        {
            std::lock_guard<std::mutex> lk(g_mu);
            m2_ctx* c = context_for(device);
            int st = m2_basis_set(c, M2_ORBITAL, h.nsh, h.l.data(), h.nprim.data(), h.exps.data(), h.coefs.data(),
                                  h.centers.data(), h.atoms.data(), pure ? 1 : 0);
            if (st == 0)
                st = m2_grad(c, h.natom, h.Z.data(), h.xyz.data(), D.data_ptr<double>(), W.data_ptr<double>(),
                             precision, g.data());
            TORCH_CHECK(st == 0, "m2.rhf_energy backward: ", m2_last_error(c));
        }
        Tensor gt = torch::from_blob(g.data(), {h.natom, 3}, torch::TensorOptions().dtype(torch::kFloat64))
                        .clone()
                        .to(R.device(), R.scalar_type());
        variable_list out(21);
        out[0] = grad_out[0].to(gt.scalar_type()) * gt;
        return out;
    }
};

Tensor rhf_energy(const Tensor& R, const Tensor& Z, const Tensor& l, const Tensor& nprim, const Tensor& exps,
                  const Tensor& coefs, const Tensor& atoms, int64_t pure, int64_t charge, int64_t multiplicity,
                  double precision, double enconverge, double densconverge, int64_t maxiter, int64_t mixed,
                  int64_t guess, const Tensor& lm, const Tensor& nprim_m, const Tensor& exps_m, const Tensor& coefs_m,
                  const Tensor& atoms_m) {
    return RHFEnergy::apply(R, Z, l, nprim, exps, coefs, atoms, pure, charge, multiplicity, precision, enconverge,
                            densconverge, maxiter, mixed, guess, lm, nprim_m, exps_m, coefs_m, atoms_m);
}

}  // namespace

TORCH_LIBRARY(m2, m) {
    m.def(
        "rhf_energy(Tensor R, Tensor Z, Tensor l, Tensor nprim, Tensor exps, Tensor coefs, Tensor atoms, int pure, "
        "int charge, int multiplicity, float precision, float enconverge, float densconverge, int maxiter, "
        "int mixed, int guess, Tensor lm, Tensor nprim_m, Tensor exps_m, Tensor coefs_m, Tensor atoms_m) -> Tensor",
        &rhf_energy);
}
