// vlm_solvers.cpp -- Anderson acceleration over the Backend's BLAS and LSQ
// (reference: vlm/src/vlm_solvers.cpp:6-150).  Every vector operation is a device BLAS call; the
// host only drives the iteration.
#include "vlm_solvers.hpp"

#include <algorithm>
#include <cstdio>

namespace vlm {
namespace {

template <typename T>
i32 anderson(Backend* backend, const TensorView<T, 1, Location::Device>& x0,
             const std::function<void(const TensorView<T, 1, Location::Device>&,
                                      const TensorView<T, 1, Location::Device>&)>& f,
             i32 max_iter, T tol_res, i32 m) {
    if (m < 1) throw std::invalid_argument("anderson_acceleration: history m must be >= 1");
    Memory* mem = backend->memory.get();
    BLAS& blas = *backend->blas;
    const i64 n = x0.shape(0);
    Tensor<T, 2, Location::Device> X{mem}, G{mem}, Gk{mem};  // Gk: LSQ input (overwritten)
    Tensor<T, 1, Location::Device> xc{mem}, xn{mem}, gc{mem}, gn{mem}, rhs{mem};
    X.init({n, m});
    G.init({n, m});
    Gk.init({n, m});
    xc.init({n});
    xn.init({n});
    gc.init({n});
    gn.init({n});
    rhs.init({n});
    auto lsq = backend->create_lsq_solver();
    lsq->init(Gk.view(), rhs.view().reshape(n, 1));
    X.view().fill(0);
    G.view().fill(0);

    // residual of one application of f at x: r = f(x) - x, with f(x) left in xn
    auto residual = [&](const TensorView<T, 1, Location::Device>& x, const TensorView<T, 1, Location::Device>& r) {
        f(x, xn.view());
        xn.view().to(r);
        blas.axpy((T)-1, x, r);
    };

    x0.to(xc.view());
    residual(xc.view(), gc.view());
    gc.view().to(X.view().slice(All, 0));
    xn.view().to(xc.view());
    residual(xc.view(), gn.view());
    gn.view().to(G.view().slice(All, 0));
    blas.axpy((T)-1, gc.view(), G.view().slice(All, 0));
    gn.view().to(gc.view());

    i32 k = 1;
    while (k < max_iter && blas.norm(gc.view()) > tol_res) {
        const i64 mk = std::min<i64>(m, k);
        const auto Gm = G.view().slice(All, Range{0, mk});
        const auto Xm = X.view().slice(All, Range{0, mk});
        const auto Gkm = Gk.view().slice(All, Range{0, mk});
        gc.view().to(rhs.view());
        Gm.to(Gkm);
        lsq->solve(Gkm, rhs.view().reshape(n, 1));
        const auto gamma = rhs.view().slice(Range{0, mk});
        // x+ = x + g - X gamma - G gamma
        xc.view().to(xn.view());
        blas.axpy((T)1, gc.view(), xn.view());
        blas.gemv((T)-1, Xm, gamma, (T)1, xn.view());
        blas.gemv((T)-1, Gm, gamma, (T)1, xn.view());
        // new history column k mod m: dx and dg
        const auto Xc = X.view().slice(All, k % m);
        const auto Gc = G.view().slice(All, k % m);
        xn.view().to(Xc);
        blas.axpy((T)-1, xc.view(), Xc);
        xn.view().to(xc.view());
        residual(xc.view(), gn.view());
        gn.view().to(Gc);
        blas.axpy((T)-1, gc.view(), Gc);
        gn.view().to(gc.view());
        k++;
    }
    xc.view().to(x0);
    if (k == max_iter) std::printf("Anderson method failed to converge\n");
    return k;
}

}  // namespace

i32 anderson_acceleration(Backend* backend, const TensorView1dD& x0,
                          const std::function<void(const TensorView1dD&, const TensorView1dD&)>& f, i32 max_iter,
                          f64 tol_res, i32 m) {
    return anderson<f64>(backend, x0, f, max_iter, tol_res, m);
}

i32 anderson_acceleration(Backend* backend, const TensorView1fD& x0,
                          const std::function<void(const TensorView1fD&, const TensorView1fD&)>& f, i32 max_iter,
                          f32 tol_res, i32 m) {
    return anderson<f32>(backend, x0, f, max_iter, tol_res, m);
}

}  // namespace vlm
