// Drives the C++ front end (include/blast/hip/Blast.hpp) on the GPU the way a
// BLAST user would call BLAST: same function names, argument order and
// exceptions.  The checks are plain loops in this file.  Prints "ALL OK".
#include <blast/hip/Blast.hpp>
#include <blast/hip/Blasfeo.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

static int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

template <class MT>
void fill_dyadic(MT& m, std::mt19937& g) {
    std::uniform_int_distribution<int> d(-256, 256);
    for (std::size_t j = 0; j < m.columns(); ++j)
        for (std::size_t i = 0; i < m.rows(); ++i) m(i, j) = d(g) / 256.0;
}

template <class MT>
void make_spd(MT& a, std::mt19937& g) {
    std::size_t const n = a.rows();
    std::normal_distribution<double> d;
    std::vector<double> gm(n * n);
    for (auto& v : gm) v = d(g);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            double s = i == j ? double(n) : 0.0;
            for (std::size_t k = 0; k < n; ++k) s += gm[i * n + k] * gm[j * n + k];
            a(i, j) = s;
        }
}

int main() {
    std::mt19937 g(42);
    // gemm NT on static panel matrices, exact dyadic data: bit-exact check.
    {
        blast::StaticPanelMatrix<double, 12, 12> A, B, C, D;
        fill_dyadic(A, g), fill_dyadic(B, g), fill_dyadic(C, g);
        blast::gemm(1.5, A, blast::trans(B), -0.5, C, D);
        for (std::size_t i = 0; i < 12; ++i)
            for (std::size_t j = 0; j < 12; ++j) {
                double s = 0.0;
                for (std::size_t k = 0; k < 12; ++k) s += A(i, k) * B(j, k);
                CHECK(D(i, j) == 1.5 * s - 0.5 * C(i, j));
            }
        blast::StaticPanelMatrix<double, 12, 11> Bad;
        bool thrown = false;
        try {
            blast::gemm(1.0, A, Bad, 1.0, C, D);
        } catch (std::invalid_argument const&) {
            thrown = true;
        }
        CHECK(thrown);
    }
    // potrf + the two trsm of a Cholesky solve on dynamic panel/dense matrices.
    {
        std::size_t const n = 20, r = 3;
        blast::DynamicPanelMatrix<double> A(n, n), L(n, n);
        blast::DynamicMatrix<double> B(n, r), X(n, r);
        make_spd(A, g);
        fill_dyadic(B, g);
        blast::potrf(A, L);
        double err = 0.0, na = 0.0;
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j <= i; ++j) {
                double s = 0.0;
                for (std::size_t k = 0; k <= j; ++k) s += L(i, k) * L(j, k);
                err = std::fmax(err, std::fabs(s - A(i, j)));
                na = std::fmax(na, std::fabs(A(i, j)));
            }
        CHECK(err <= 1e-13 * na);
        blast::trsm<blast::UpLo::Lower, false>(L, B, X);
        blast::trsm<blast::UpLo::Upper, false>(blast::trans(L), X, X);
        double res = 0.0;
        for (std::size_t j = 0; j < r; ++j)
            for (std::size_t i = 0; i < n; ++i) {
                double s = 0.0;
                for (std::size_t k = 0; k < n; ++k) s += A(i, k) * X(k, j);
                res = std::fmax(res, std::fabs(s - B(i, j)));
            }
        CHECK(res <= 1e-12);
        bool thrown = false;
        blast::DynamicPanelMatrix<double> Wide(4, 8), Wl(4, 8);
        try {
            blast::potrf(Wide, Wl);
        } catch (std::invalid_argument const&) {
            thrown = true;
        }
        CHECK(thrown);
    }
    // getrf with partial pivoting; singular matrix throws.
    {
        std::size_t const n = 9;
        blast::DynamicMatrix<double> A(n, n), A0(n, n);
        std::uniform_real_distribution<double> u(-1, 1);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < n; ++i) A(i, j) = A0(i, j) = u(g);
        std::size_t ipiv[9];
        blast::getrf(A, ipiv);
        std::vector<double> P(n * n);
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j) P[i * n + j] = A0(i, j);
        for (std::size_t k = 0; k < n; ++k)
            if (ipiv[k] != k)
                for (std::size_t j = 0; j < n; ++j) std::swap(P[k * n + j], P[ipiv[k] * n + j]);
        double err = 0.0;
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j) {
                double s = 0.0;
                for (std::size_t k = 0; k <= std::min(i, j); ++k) s += (k == i ? 1.0 : A(i, k)) * A(k, j);
                err = std::fmax(err, std::fabs(s - P[i * n + j]));
            }
        CHECK(err <= 1e-13);
        blast::DynamicMatrix<double> S(4, 4);
        for (std::size_t j = 0; j < 4; ++j)
            for (std::size_t i = 0; i < 4; ++i) S(i, j) = 1.0;
        bool thrown = false;
        try {
            blast::getrf(S, ipiv);
        } catch (std::invalid_argument const&) {
            thrown = true;
        }
        CHECK(thrown);
    }
    // syrkLower and the unimplemented trmm form.
    {
        blast::DynamicMatrix<double> A(7, 5), C(7, 7), D(7, 7);
        fill_dyadic(A, g), fill_dyadic(C, g);
        blast::syrkLower(0.5, A, 2.0, C, D);
        for (std::size_t j = 0; j < 7; ++j)
            for (std::size_t i = j; i < 7; ++i) {
                double s = 0.0;
                for (std::size_t k = 0; k < 5; ++k) s += A(i, k) * A(j, k);
                CHECK(D(i, j) == 0.5 * s + 2.0 * C(i, j));
            }
        bool thrown = false;
        blast::hip::DeviceArray<double> buf(64);
        blast::StaticMatrix<double, 4, 4> proto;
        auto b = blast::hip::Batch<double>::like(proto, buf.get(), 1);
        try {
            blast::trmm<double>(1.0, b, blast::UpLo::Upper, true, b, b);
        } catch (std::logic_error const&) {
            thrown = true;
        }
        CHECK(thrown);
    }
    // Batched Cholesky solve on device memory: 64 instances of n = 32.
    {
        std::size_t const n = 32, r = 2, batch = 64;
        blast::DynamicPanelMatrix<double> protoA(n, n), protoB(n, r);
        std::size_t const sa = blast::nextMultiple(protoA.storageSize(), 8), sb = blast::nextMultiple(protoB.storageSize(), 8);
        std::vector<double> ha(batch * sa), hb(batch * sb), hx(batch * sb);
        std::vector<blast::DynamicPanelMatrix<double>> As;
        for (std::size_t k = 0; k < batch; ++k) {
            blast::DynamicPanelMatrix<double> a(n, n), bm(n, r);
            make_spd(a, g);
            fill_dyadic(bm, g);
            std::memcpy(&ha[k * sa], a.data(), a.storageSize() * 8);
            std::memcpy(&hb[k * sb], bm.data(), bm.storageSize() * 8);
        }
        blast::hip::DeviceArray<double> da(ha.size()), dl(ha.size()), db(hb.size()), dx(hb.size());
        blast::hip::DeviceArray<std::int32_t> info(batch);
        blast::hip::check_hip(hipMemcpy(da.get(), ha.data(), ha.size() * 8, hipMemcpyHostToDevice));
        blast::hip::check_hip(hipMemcpy(db.get(), hb.data(), hb.size() * 8, hipMemcpyHostToDevice));
        auto A = blast::hip::Batch<double>::like(protoA, da.get(), batch);
        auto L = blast::hip::Batch<double>::like(protoA, dl.get(), batch);
        auto B = blast::hip::Batch<double>::like(protoB, db.get(), batch);
        auto X = blast::hip::Batch<double>::like(protoB, dx.get(), batch);
        blast::posv(A, L, B, X, info.get());
        blast::hip::check_hip(hipDeviceSynchronize());
        blast::hip::check_hip(hipMemcpy(hx.data(), dx.get(), hx.size() * 8, hipMemcpyDeviceToHost));
        double worst = 0.0;
        for (std::size_t k = 0; k < batch; ++k) {
            blast::DynamicPanelMatrix<double> a(n, n), bm(n, r), xm(n, r);
            std::memcpy(a.data(), &ha[k * sa], a.storageSize() * 8);
            std::memcpy(bm.data(), &hb[k * sb], bm.storageSize() * 8);
            std::memcpy(xm.data(), &hx[k * sb], xm.storageSize() * 8);
            for (std::size_t j = 0; j < r; ++j)
                for (std::size_t i = 0; i < n; ++i) {
                    double s = 0.0;
                    for (std::size_t q = 0; q < n; ++q) s += (i >= q ? a(i, q) : a(q, i)) * xm(q, j);
                    worst = std::fmax(worst, std::fabs(s - bm(i, j)));
                }
        }
        CHECK(worst <= 1e-12);
    }
    // Storage conversion on the device: panel -> row-major dense, bit-exact.
    {
        std::size_t const m = 37, n = 21, batch = 5;
        blast::DynamicPanelMatrix<double> P(m, n);
        blast::DynamicMatrix<double, blast::rowMajor> R(m, n);
        std::size_t const sp = blast::nextMultiple(P.storageSize(), 8), sr = blast::nextMultiple(R.storageSize(), 8);
        std::vector<double> hp(batch * sp), hr(batch * sr, -1.0);
        for (std::size_t k = 0; k < batch; ++k) {
            fill_dyadic(P, g);
            std::memcpy(&hp[k * sp], P.data(), P.storageSize() * 8);
        }
        blast::hip::DeviceArray<double> dp(hp.size()), dr(hr.size());
        blast::hip::check_hip(hipMemcpy(dp.get(), hp.data(), hp.size() * 8, hipMemcpyHostToDevice));
        auto S = blast::hip::Batch<double>::like(P, dp.get(), batch);
        auto D = blast::hip::Batch<double>::like(R, dr.get(), batch);
        blast::assign<double>(S, D);
        blast::hip::check_hip(hipMemcpy(hr.data(), dr.get(), hr.size() * 8, hipMemcpyDeviceToHost));
        for (std::size_t k = 0; k < batch; ++k) {
            std::memcpy(P.data(), &hp[k * sp], P.storageSize() * 8);
            std::memcpy(R.data(), &hr[k * sr], R.storageSize() * 8);
            for (std::size_t i = 0; i < m; ++i)
                for (std::size_t j = 0; j < n; ++j) CHECK(R(i, j) == P(i, j));
        }
    }
    // Batched LU solve (gesv) on 16 instances of n = 40.
    {
        std::size_t const n = 40, batch = 16;
        blast::DynamicMatrix<double> protoA(n, n), protoB(n, 1);
        std::size_t const sa = blast::nextMultiple(protoA.storageSize(), 8), sb = blast::nextMultiple(protoB.storageSize(), 8);
        std::vector<double> ha(batch * sa), hb(batch * sb), hx(batch * sb);
        std::uniform_real_distribution<double> u(-1, 1);
        for (auto& v : ha) v = u(g);
        for (auto& v : hb) v = u(g);
        for (std::size_t k = 0; k < batch; ++k)
            for (std::size_t i = 0; i < n; ++i) ha[k * sa + i + i * n] += 4.0;  // well conditioned
        std::vector<double> ha0 = ha;
        blast::hip::DeviceArray<double> da(ha.size()), db(hb.size()), dx(hb.size());
        blast::hip::DeviceArray<std::int64_t> piv(batch * n);
        blast::hip::check_hip(hipMemcpy(da.get(), ha.data(), ha.size() * 8, hipMemcpyHostToDevice));
        blast::hip::check_hip(hipMemcpy(db.get(), hb.data(), hb.size() * 8, hipMemcpyHostToDevice));
        auto A = blast::hip::Batch<double>::like(protoA, da.get(), batch);
        auto B = blast::hip::Batch<double>::like(protoB, db.get(), batch);
        auto X = blast::hip::Batch<double>::like(protoB, dx.get(), batch);
        blast::gesv(A, piv.get(), std::int64_t(n), B, X);
        blast::hip::check_hip(hipMemcpy(hx.data(), dx.get(), hx.size() * 8, hipMemcpyDeviceToHost));
        double worst = 0.0;
        for (std::size_t k = 0; k < batch; ++k)
            for (std::size_t i = 0; i < n; ++i) {
                double s = 0.0;
                for (std::size_t q = 0; q < n; ++q) s += ha0[k * sa + i + q * n] * hx[k * sb + q];
                worst = std::fmax(worst, std::fabs(s - hb[k * sb + i]));
            }
        CHECK(worst <= 1e-11);
    }
    // Uneven Riccati batch: scalar instances (nx = nu = 1) checked against the
    // closed form of the scalar recursion.
    {
        int const horizon = 5;
        std::vector<std::int32_t> nx = {1, 1, 3}, nu = {1, 1, 2};
        blast::hip::RiccatiBatch rb(nx, nu, horizon);
        std::vector<double> in(rb.inputSize(), 0.0);
        double const av[2] = {1.3, 0.6}, bv[2] = {0.7, 1.1}, qv[2] = {2.0, 1.0}, rv[2] = {0.5, 3.0};
        auto const& ins = rb.instances();
        for (int k = 0; k < 2; ++k) {
            in[ins[k].a] = av[k], in[ins[k].b] = bv[k], in[ins[k].q] = qv[k], in[ins[k].r] = rv[k];
        }
        for (int i = 0; i < 3; ++i) in[ins[2].q + i * 3 + i] = 1.0, in[ins[2].a + i * 3 + i] = 0.5;  // Q = I, A = 0.5 I
        for (int i = 0; i < 2; ++i) in[ins[2].r + i * 2 + i] = 1.0, in[ins[2].b + i * 3 + i] = 1.0;
        rb.setInputs(in);
        rb.run();
        blast::hip::check_hip(hipDeviceSynchronize());
        auto out = rb.outputs();
        auto info = rb.info();
        for (int k = 0; k < 2; ++k) {
            double p = qv[k];
            double kk = 0.0;
            for (int s = horizon - 1; s >= 0; --s) {
                kk = -av[k] * bv[k] * p / (rv[k] + bv[k] * bv[k] * p);
                CHECK(std::fabs(out[ins[k].k + s] - kk) <= 1e-14 * std::fabs(kk));
                p = qv[k] + av[k] * av[k] * p * rv[k] / (rv[k] + bv[k] * bv[k] * p);
            }
            CHECK(std::fabs(out[ins[k].l] * out[ins[k].l] - p) <= 1e-14 * p);
            CHECK(info[k] == 0);
        }
        CHECK(info[2] == 0);
    }
    // BLASFEO-format wrappers (blast/hip/Blasfeo.hpp): gemm_nt on sub-blocks
    // with exact dyadic data (bit-exact), syrk_potrf, and an argument error.
    {
        namespace bf = blast::hip::blasfeo;
        using BB = bf::BlasfeoBatch<double>;
        int const batch = 3, m = 10, n = 9, k = 7;
        std::int64_t const sa = BB::storage(4 + m, k + 3), sb = BB::storage(n, k + 1), sc = BB::storage(m, n),
                           sd = BB::storage(m + 4, n + 2);
        std::vector<double> ha(batch * sa, 0.0), hb(batch * sb, 0.0), hc(batch * sc, 0.0), hd(batch * sd, 9.0);
        BB A0(nullptr, 4 + m, k + 3, k + 3, sa, batch), B0(nullptr, n, k + 1, k + 1, sb, batch),
            C0(nullptr, m, n, n, sc, batch), D0(nullptr, m + 4, n + 2, n + 2, sd, batch);
        std::uniform_int_distribution<int> dd(-64, 64);
        for (auto* v : {&ha, &hb, &hc})
            for (auto& x : *v) x = dd(g) / 64.0;
        blast::hip::DeviceArray<double> da(ha.size()), db(hb.size()), dc(hc.size()), dv(hd.size());
        blast::hip::check_hip(hipMemcpy(da.get(), ha.data(), ha.size() * 8, hipMemcpyHostToDevice));
        blast::hip::check_hip(hipMemcpy(db.get(), hb.data(), hb.size() * 8, hipMemcpyHostToDevice));
        blast::hip::check_hip(hipMemcpy(dc.get(), hc.data(), hc.size() * 8, hipMemcpyHostToDevice));
        blast::hip::check_hip(hipMemcpy(dv.get(), hd.data(), hd.size() * 8, hipMemcpyHostToDevice));
        BB A = A0, B = B0, C = C0, D = D0;
        A.pA = da.get(), B.pA = db.get(), C.pA = dc.get(), D.pA = dv.get();
        // D[4.., 2..] = 0.5*C + 1.5 * A[4.., 3..] * B[0.., 1..]^T
        bf::gemm_nt(m, n, k, 1.5, A, 4, 3, B, 0, 1, 0.5, C, 0, 0, D, 4, 2);
        blast::hip::check_hip(hipDeviceSynchronize());
        std::vector<double> out(hd.size());
        blast::hip::check_hip(hipMemcpy(out.data(), dv.get(), out.size() * 8, hipMemcpyDeviceToHost));
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < m + 4; ++i)
                for (int j = 0; j < n + 2; ++j) {
                    double const got = out[b * sd + D0.index(i, j)];
                    if (i < 4 || j < 2) {
                        CHECK(got == 9.0);
                        continue;
                    }
                    double s = 0.0;
                    for (int l = 0; l < k; ++l) s += ha[b * sa + A0.index(i, 3 + l)] * hb[b * sb + B0.index(j - 2, 1 + l)];
                    CHECK(got == 0.5 * hc[b * sc + C0.index(i - 4, j - 2)] + 1.5 * s);
                }
        // syrk_potrf: D = chol(S + A A^T) with S = m I, A = the first k columns of A's rows 4..
        int const mm = 8, kk = 5;
        std::int64_t const ss = BB::storage(mm, mm);
        std::vector<double> hs(batch * ss, 0.0);
        BB S0(nullptr, mm, mm, mm, ss, batch);
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < mm; ++i) hs[b * ss + S0.index(i, i)] = double(mm);
        blast::hip::DeviceArray<double> ds(hs.size()), dl(hs.size());
        blast::hip::check_hip(hipMemcpy(ds.get(), hs.data(), hs.size() * 8, hipMemcpyHostToDevice));
        BB S = S0, Lf = S0;
        S.pA = ds.get(), Lf.pA = dl.get();
        bf::syrk_potrf(mm, kk, A, 4, 0, A, 4, 0, S, 0, 0, Lf, 0, 0);
        blast::hip::check_hip(hipDeviceSynchronize());
        std::vector<double> hl(hs.size());
        blast::hip::check_hip(hipMemcpy(hl.data(), dl.get(), hl.size() * 8, hipMemcpyDeviceToHost));
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < mm; ++i)
                for (int j = 0; j <= i; ++j) {
                    double llt = 0.0, t = i == j ? double(mm) : 0.0;
                    for (int l = 0; l <= j; ++l) llt += hl[b * ss + S0.index(i, l)] * hl[b * ss + S0.index(j, l)];
                    for (int l = 0; l < kk; ++l) t += ha[b * sa + A0.index(4 + i, l)] * ha[b * sa + A0.index(4 + j, l)];
                    CHECK(std::fabs(llt - t) <= 1e-13 * mm);
                }
        bool thrown = false;
        try {
            bf::potrf(mm, S, 2, 0, Lf, 0, 0);  // row offset 2 splits a panel
        } catch (std::invalid_argument const&) {
            thrown = true;
        }
        CHECK(thrown);
    }
    // Views: gemm on submatrices at offsets (dense, and panel with a row offset
    // on a panel boundary), the pointer form, rows of a panel matrix (inside a
    // panel: elements panel-size apart), and the panel alignment rule.
    {
        std::mt19937 g(11);
        blast::DynamicPanelMatrix<double> A(40, 30), Cp(36, 28), Dp(36, 28);
        blast::DynamicMatrix<double> B(30, 25);
        fill_dyadic(A, g), fill_dyadic(B, g), fill_dyadic(Cp, g), fill_dyadic(Dp, g);
        blast::DynamicPanelMatrix<double> D0(36, 28);
        for (std::size_t j = 0; j < 28; ++j)
            for (std::size_t i = 0; i < 36; ++i) D0(i, j) = Dp(i, j);
        // D[8..8+20, 3..3+11] = 1.5 A[4..24, 5..14] B[2..11, 7..18] - 0.5 C[8.., 3..]
        auto d = blast::submatrix(Dp, 8, 3, 20, 11);
        blast::gemm(1.5, blast::submatrix(A, 4, 5, 20, 9), blast::submatrix(B, 2, 7, 9, 11), -0.5,
                    blast::submatrix(Cp, 8, 3, 20, 11), d);
        bool ok = true;
        for (std::size_t j = 0; j < 28; ++j)
            for (std::size_t i = 0; i < 36; ++i) {
                double ref = D0(i, j);
                if (i >= 8 && i < 28 && j >= 3 && j < 14) {
                    double acc = 0.0;
                    for (std::size_t k = 0; k < 9; ++k) acc += A(4 + i - 8, 5 + k) * B(2 + k, 7 + j - 3);
                    ref = 1.5 * acc - 0.5 * Cp(i, j);
                }
                ok = ok && Dp(i, j) == ref;
            }
        CHECK(ok);
        // the same product through the pointer form, into a fresh D
        blast::DynamicPanelMatrix<double> D2(36, 28);
        blast::gemm(20, 11, 9, 1.5, blast::ptr(A, 4, 5), blast::ptr(B, 2, 7), -0.5, blast::ptr(Cp, 8, 3),
                    blast::ptr(D2, 8, 3));
        ok = true;
        for (std::size_t j = 3; j < 14; ++j)
            for (std::size_t i = 8; i < 28; ++i) ok = ok && D2(i, j) == Dp(i, j);
        CHECK(ok);
        // row 5 of A (inside a panel) times B[0..30, 0..25] into row 9 of a dense D
        blast::DynamicMatrix<double> Dr(12, 25), Cr(12, 25);
        fill_dyadic(Cr, g);
        auto dr = blast::row(Dr, 9);
        blast::gemm(1.0, blast::row(A, 5), B, 1.0, blast::row(Cr, 9), dr);
        ok = true;
        for (std::size_t j = 0; j < 25; ++j) {
            double acc = Cr(9, j);
            for (std::size_t k = 0; k < 30; ++k) acc += A(5, k) * B(k, j);
            ok = ok && Dr(9, j) == acc && Dr(8, j) == 0.0;
        }
        CHECK(ok);
        bool threw = false;
        try {
            (void)blast::submatrix(A, 2, 0, 4, 4);  // row 2 splits a panel
        } catch (std::logic_error const&) {
            threw = true;
        }
        CHECK(threw);
        threw = false;
        try {
            (void)blast::submatrix(A, 36, 0, 8, 4);
        } catch (std::invalid_argument const&) {
            threw = true;
        }
        CHECK(threw);
    }

    // trmm, both matrix forms (algorithm/Trmm.hpp:260-310), and the
    // unimplemented triangles.
    {
        std::mt19937 g(12);
        blast::DynamicPanelMatrix<double> U(16, 16), Lo(12, 12), B1(16, 12), C1(16, 12), B2(16, 12), C2(16, 12);
        fill_dyadic(U, g), fill_dyadic(Lo, g), fill_dyadic(B1, g), fill_dyadic(B2, g);
        blast::trmm(2.0, U, blast::UpLo::Upper, false, B1, C1);
        blast::trmm(0.5, B2, Lo, blast::UpLo::Lower, false, C2);
        bool ok = true;
        for (std::size_t i = 0; i < 16; ++i)
            for (std::size_t j = 0; j < 12; ++j) {
                double a1 = 0.0, a2 = 0.0;
                for (std::size_t k = i; k < 16; ++k) a1 += U(i, k) * B1(k, j);
                for (std::size_t k = j; k < 12; ++k) a2 += B2(i, k) * Lo(k, j);
                ok = ok && C1(i, j) == 2.0 * a1 && C2(i, j) == 0.5 * a2;
            }
        CHECK(ok);
        bool threw = false;
        try {
            blast::trmm(1.0, U, blast::UpLo::Lower, false, B1, C1);
        } catch (std::logic_error const&) {
            threw = true;
        }
        CHECK(threw);
    }

    // trsvLeftLower / trsvLeftUpper (dense/Trsv.hpp:15-68): banded dyadic
    // triangles with unit-magnitude diagonals and integer x0, so b = A x0 is
    // exact; the solve (inverted 16x16 diagonal tiles on MFMA) returns x0 to
    // rounding.
    {
        std::mt19937 g(13);
        std::uniform_int_distribution<int> d(-3, 3);
        for (std::size_t n : {1u, 7u, 40u, 129u}) {
            blast::DynamicMatrix<double> Al(n, n), Au(n, n);
            std::vector<double> x0(n), bl(n, 0.0), bu(n, 0.0), xl, xu;
            for (std::size_t i = 0; i < n; ++i) x0[i] = d(g);
            for (std::size_t i = 0; i < n; ++i)
                for (std::size_t j = 0; j < n; ++j) {
                    double const v = i == j ? ((i % 2) ? -1.0 : 1.0) : (std::abs(int(i) - int(j)) < 3 ? d(g) / 8.0 : 0.0);
                    Al(i, j) = j <= i ? v : 0.0;
                    Au(i, j) = j >= i ? v : 0.0;
                }
            for (std::size_t i = 0; i < n; ++i)
                for (std::size_t j = 0; j < n; ++j) bl[i] += Al(i, j) * x0[j], bu[i] += Au(i, j) * x0[j];
            blast::trsvLeftLower(Al, bl, xl);
            blast::trsvLeftUpper(Au, bu, xu);
            double el = 0.0, eu = 0.0;
            for (std::size_t i = 0; i < n; ++i) el = std::max(el, std::abs(xl[i] - x0[i])), eu = std::max(eu, std::abs(xu[i] - x0[i]));
            CHECK(xl.size() == n && el < 1e-12 * double(n));
            CHECK(xu.size() == n && eu < 1e-12 * double(n));
        }
        bool threw = false;
        try {
            std::vector<double> b(3), x;
            blast::trsvLeftLower(blast::DynamicMatrix<double>(4, 4), b, x);
        } catch (std::invalid_argument const&) {
            threw = true;
        }
        CHECK(threw);
    }

    if (failures == 0) std::printf("ALL OK\n");
    return failures ? 1 : 0;
}
