// CPU ORACLE -- test infrastructure only. The learned seeder's network (the
// engine's csrc/seeder.hpp: four 3x3 stride-2 convolutions with ReLU, 2 -> 32
// -> 32 -> 32 -> 32 channels on a 2 x 64 x 64 input, then a 512 -> 6 linear
// layer, mean squared error over B x 6) restated in double precision, forward
// and backward, as the checker of the GPU kernels' gradients. It has no
// reference counterpart; it is pinned by central finite differences of its
// own loss (tests/test_seeder_oracle.py).
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr int kCh = 32, kIn = 2, kP = 64, kFeat = 512, kOut = 6;
int conv_w(int l) { return kCh * (l == 0 ? kIn : kCh) * 9; }
int off_w(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += conv_w(k) + kCh;
    return o;
}
int off_b(int l) { return off_w(l) + conv_w(l); }
int off_fw() { return off_b(3) + kCh; }
int off_fb() { return off_fw() + kOut * kFeat; }
int total() { return off_fb() + kOut; }

struct Net {
    int B;
    std::vector<double> x, a[4], y;   // a[l]: B x 32 x h x h after ReLU
};

void conv_fwd(const double* p, int l, const std::vector<double>& in, int cin, int hin, std::vector<double>& out,
              int B) {
    const int ho = hin / 2;
    out.assign((size_t)B * kCh * ho * ho, 0.0);
    const double* w = p + off_w(l);
    const double* b = p + off_b(l);
    for (int n = 0; n < B; ++n)
        for (int co = 0; co < kCh; ++co)
            for (int y = 0; y < ho; ++y)
                for (int x = 0; x < ho; ++x) {
                    double s = b[co];
                    for (int ci = 0; ci < cin; ++ci)
                        for (int ky = 0; ky < 3; ++ky) {
                            const int iy = 2 * y + ky - 1;
                            if (iy < 0 || iy >= hin) continue;
                            for (int kx = 0; kx < 3; ++kx) {
                                const int ix = 2 * x + kx - 1;
                                if (ix < 0 || ix >= hin) continue;
                                s += w[((co * cin + ci) * 3 + ky) * 3 + kx] *
                                     in[(((size_t)n * cin + ci) * hin + iy) * hin + ix];
                            }
                        }
                    out[(((size_t)n * kCh + co) * ho + y) * ho + x] = s > 0.0 ? s : 0.0;
                }
}

void forward(const double* p, Net& net) {
    conv_fwd(p, 0, net.x, kIn, kP, net.a[0], net.B);
    conv_fwd(p, 1, net.a[0], kCh, 32, net.a[1], net.B);
    conv_fwd(p, 2, net.a[1], kCh, 16, net.a[2], net.B);
    conv_fwd(p, 3, net.a[2], kCh, 8, net.a[3], net.B);
    net.y.assign((size_t)net.B * kOut, 0.0);
    for (int n = 0; n < net.B; ++n)
        for (int k = 0; k < kOut; ++k) {
            double s = p[off_fb() + k];
            for (int j = 0; j < kFeat; ++j) s += p[off_fw() + k * kFeat + j] * net.a[3][(size_t)n * kFeat + j];
            net.y[(size_t)n * kOut + k] = s;
        }
}

double loss_of(const Net& net, const float* t) {
    double l = 0.0;
    for (size_t i = 0; i < net.y.size(); ++i) {
        const double d = net.y[i] - t[i];
        l += d * d;
    }
    return l / (double)net.y.size();
}

// g: gradient of the loss w.r.t. the pre-activation output of layer l (B x 32
// x ho x ho); accumulates the weight and bias gradients and, for l > 0, the
// gradient w.r.t. the layer's input, masked by that input's ReLU.
void conv_bwd(const double* p, int l, const std::vector<double>& in, int cin, int hin, const std::vector<double>& g,
              double* gp, std::vector<double>* gin, int B) {
    const int ho = hin / 2;
    const double* w = p + off_w(l);
    double* gw = gp + off_w(l);
    double* gb = gp + off_b(l);
    if (gin) gin->assign((size_t)B * cin * hin * hin, 0.0);
    for (int n = 0; n < B; ++n)
        for (int co = 0; co < kCh; ++co)
            for (int y = 0; y < ho; ++y)
                for (int x = 0; x < ho; ++x) {
                    const double gv = g[(((size_t)n * kCh + co) * ho + y) * ho + x];
                    gb[co] += gv;
                    for (int ci = 0; ci < cin; ++ci)
                        for (int ky = 0; ky < 3; ++ky) {
                            const int iy = 2 * y + ky - 1;
                            if (iy < 0 || iy >= hin) continue;
                            for (int kx = 0; kx < 3; ++kx) {
                                const int ix = 2 * x + kx - 1;
                                if (ix < 0 || ix >= hin) continue;
                                const size_t ii = (((size_t)n * cin + ci) * hin + iy) * hin + ix;
                                const int wi = ((co * cin + ci) * 3 + ky) * 3 + kx;
                                gw[wi] += gv * in[ii];
                                if (gin) (*gin)[ii] += gv * w[wi];
                            }
                        }
                }
    if (gin)
        for (size_t i = 0; i < gin->size(); ++i)
            if (!(in[i] > 0.0)) (*gin)[i] = 0.0;
}

}  // namespace

extern "C" {

int orc_seeder_param_count(void) { return total(); }

// y: B x 6 outputs; grads (optional): gradient of the mean squared error
// against t; returns the loss (0 when t is null).
double orc_seeder_pass(const float* params, const float* x, const float* t, int B, float* y, float* grads) {
    std::vector<double> p(params, params + total());
    Net net;
    net.B = B;
    net.x.assign(x, x + (size_t)B * kIn * kP * kP);
    forward(p.data(), net);
    if (y)
        for (size_t i = 0; i < net.y.size(); ++i) y[i] = (float)net.y[i];
    if (!t) return 0.0;
    const double loss = loss_of(net, t);
    if (grads) {
        std::vector<double> gp(total(), 0.0), gy(net.y.size());
        for (size_t i = 0; i < gy.size(); ++i) gy[i] = 2.0 * (net.y[i] - t[i]) / (double)gy.size();
        std::vector<double> g4((size_t)B * kFeat, 0.0);
        for (int n = 0; n < B; ++n)
            for (int k = 0; k < kOut; ++k) {
                const double gv = gy[(size_t)n * kOut + k];
                gp[off_fb() + k] += gv;
                for (int j = 0; j < kFeat; ++j) {
                    gp[off_fw() + k * kFeat + j] += gv * net.a[3][(size_t)n * kFeat + j];
                    g4[(size_t)n * kFeat + j] += gv * p[off_fw() + k * kFeat + j];
                }
            }
        for (size_t i = 0; i < g4.size(); ++i)
            if (!(net.a[3][i] > 0.0)) g4[i] = 0.0;
        std::vector<double> g3, g2, g1;
        conv_bwd(p.data(), 3, net.a[2], kCh, 8, g4, gp.data(), &g3, B);
        conv_bwd(p.data(), 2, net.a[1], kCh, 16, g3, gp.data(), &g2, B);
        conv_bwd(p.data(), 1, net.a[0], kCh, 32, g2, gp.data(), &g1, B);
        conv_bwd(p.data(), 0, net.x, kIn, 64, g1, gp.data(), nullptr, B);
        for (int i = 0; i < total(); ++i) grads[i] = (float)gp[i];
    }
    return loss;
}

// The loss at params + h e_i and params - h e_i (double throughout), for the
// finite-difference check of the analytic gradient.
double orc_seeder_loss_at(const float* params, const float* x, const float* t, int B, int i, double h) {
    std::vector<double> p(params, params + total());
    p[i] += h;
    Net net;
    net.B = B;
    net.x.assign(x, x + (size_t)B * kIn * kP * kP);
    forward(p.data(), net);
    return loss_of(net, t);
}

}  // extern "C"
