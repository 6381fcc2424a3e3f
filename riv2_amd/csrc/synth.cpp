// Seeded synthetic RGB24 sequence generator for the benchmark and parity workloads.
//
//  * background luma: periodic Gaussian field with a 1/f^2 power spectrum (inverse FFT of
//    shaped complex white noise), min/max scaled to [16, 235];
//  * chroma: smoother periodic fields (1/f^4 power), scaled to [96, 160];
//  * global translation per frame drawn uniformly in [-M, M] pixels at half-pel steps;
//  * K textured rectangles (own 1/f^2 texture, own flat chroma) moving with an integer
//    velocity drawn in [-V, V] px/frame, reflected at the frame border; optionally each zooms
//    and rotates about its centre at its own rate (drawn in [-Z, Z] per frame, [-R, R] degrees
//    per frame);
//  * optional hard scene cut every C frames (new background phase and chroma);
//  * additive Gaussian noise (sigma) on Y, U and V, rounded and clipped;
//  * RGB24 from the YUV planes through the codec's BT.709 limited-range inverse.
//
// Everything is a function of (params, seed): frames are rendered independently from a
// precomputed trajectory, so any frame range can be regenerated byte-identically.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <complex>
#include <vector>
#include <thread>
#include <algorithm>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {   // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
    double gauss() {   // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

typedef std::complex<double> cd;
static void fft1d(cd* a, int n, bool inv) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = 2 * M_PI / len * (inv ? 1 : -1);
        cd wl(std::cos(ang), std::sin(ang));
        for (int i = 0; i < n; i += len) {
            cd w(1);
            for (int k = 0; k < len / 2; k++) {
                cd u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v; a[i + k + len / 2] = u - v;
                w *= wl;
            }
        }
    }
}
// Periodic N x N field with power spectrum ~ 1/f^beta, normalised to zero mean / unit std.
static std::vector<float> make_texture(int n, double beta, Rng& rng) {
    std::vector<cd> a((size_t)n * n);
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
            int kx = x <= n / 2 ? x : x - n, ky = y <= n / 2 ? y : y - n;
            double f = std::sqrt((double)kx * kx + (double)ky * ky);
            double amp = f == 0 ? 0.0 : std::pow(f, -beta / 2.0);
            double g1 = rng.gauss(), g2 = rng.gauss();
            a[(size_t)y * n + x] = cd(g1 * amp, g2 * amp);
        }
    std::vector<cd> tmp(n);
    for (int y = 0; y < n; y++) fft1d(&a[(size_t)y * n], n, true);
    for (int x = 0; x < n; x++) {
        for (int y = 0; y < n; y++) tmp[y] = a[(size_t)y * n + x];
        fft1d(tmp.data(), n, true);
        for (int y = 0; y < n; y++) a[(size_t)y * n + x] = tmp[y];
    }
    std::vector<float> t((size_t)n * n);
    double sum = 0, sq = 0;
    for (size_t i = 0; i < t.size(); i++) { double v = a[i].real(); sum += v; sq += v * v; }
    double mean = sum / t.size(), sd = std::sqrt(std::max(1e-30, sq / t.size() - mean * mean));
    for (size_t i = 0; i < t.size(); i++) t[i] = (float)((a[i].real() - mean) / sd);
    return t;
}
static void scale_minmax(std::vector<float>& t, float lo, float hi) {
    float mn = *std::min_element(t.begin(), t.end()), mx = *std::max_element(t.begin(), t.end());
    float s = (hi - lo) / std::max(1e-6f, mx - mn);
    for (auto& v : t) v = lo + (v - mn) * s;
}
static inline float sample_wrap(const std::vector<float>& t, int n, double x, double y) {
    double fx = std::floor(x), fy = std::floor(y);
    double ax = x - fx, ay = y - fy;
    int x0 = ((int)fx % n + n) % n, y0 = ((int)fy % n + n) % n;
    int x1 = (x0 + 1) % n, y1 = (y0 + 1) % n;
    double a = t[(size_t)y0 * n + x0], b = t[(size_t)y0 * n + x1], c = t[(size_t)y1 * n + x0], d = t[(size_t)y1 * n + x1];
    return (float)((a * (1 - ax) + b * ax) * (1 - ay) + (c * (1 - ax) + d * ax) * ay);
}
static inline uint8_t clip8(double v) { long r = std::lround(v); return (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r)); }

}  // namespace

extern "C" {

typedef struct {
    uint32_t width, height, frames;
    uint64_t seed;
    float global_motion;        // M: per-frame global translation drawn in [-M, M] px, half-pel steps
    uint32_t n_rects, rect_min, rect_max;
    float rect_velocity;        // V: per-rectangle velocity drawn in [-V, V] px/frame (integer)
    float noise_sigma;
    uint32_t scene_cut_interval;   // 0 = none
    float rect_zoom;               // Z: per-rectangle zoom rate drawn in [-Z, Z] per frame (0.005 = 0.5%)
    float rect_rotate_deg;         // R: per-rectangle rotation rate drawn in [-R, R] degrees per frame
} riv_synth_params;

struct riv_synth {
    riv_synth_params p;
    int nbg, nch, nrt;
    std::vector<float> bg, cu, cv;
    struct Rect { int w, h; std::vector<float> tex; uint8_t u, v; std::vector<int> px, py; double zoom = 0, rot = 0; };
    std::vector<Rect> rects;
    std::vector<int> gx, gy;        // cumulative global offset, half-pel units
    std::vector<int> cut_dx, cut_dy; // background phase after scene cuts (pixels)
};

int riv_synth_create(const riv_synth_params* p, riv_synth** out) {
    if (!p || p->width < 16 || p->height < 16 || p->width % 2 || p->height % 2 || p->frames == 0) return 1;
    riv_synth* s = new riv_synth();
    s->p = *p;
    Rng rng(p->seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
    int n = 256; while (n < (int)std::max(p->width, p->height)) n <<= 1;
    s->nbg = n; s->nch = std::max(64, n / 2); s->nrt = 128;
    s->bg = make_texture(n, 2.0, rng); scale_minmax(s->bg, 16.0f, 235.0f);
    s->cu = make_texture(s->nch, 4.0, rng); scale_minmax(s->cu, 96.0f, 160.0f);
    s->cv = make_texture(s->nch, 4.0, rng); scale_minmax(s->cv, 96.0f, 160.0f);
    int F = (int)p->frames, W = (int)p->width, H = (int)p->height;
    int M2 = (int)std::lround(2.0 * p->global_motion);
    s->gx.assign(F, 0); s->gy.assign(F, 0); s->cut_dx.assign(F, 0); s->cut_dy.assign(F, 0);
    int cdx = 0, cdy = 0;
    for (int f = 1; f < F; f++) {
        s->gx[f] = s->gx[f - 1] + rng.range(-M2, M2);
        s->gy[f] = s->gy[f - 1] + rng.range(-M2, M2);
        if (p->scene_cut_interval && f % (int)p->scene_cut_interval == 0) { cdx = rng.range(0, n - 1); cdy = rng.range(0, n - 1); }
        s->cut_dx[f] = cdx; s->cut_dy[f] = cdy;
    }
    int V = (int)std::lround(p->rect_velocity);
    for (uint32_t k = 0; k < p->n_rects; k++) {
        riv_synth::Rect r;
        r.w = rng.range((int)p->rect_min, (int)p->rect_max); r.h = rng.range((int)p->rect_min, (int)p->rect_max);
        r.w = std::min(r.w, W); r.h = std::min(r.h, H);
        r.tex = make_texture(s->nrt, 2.0, rng);
        float lo = (float)rng.range(16, 120), hi = (float)rng.range(130, 235);
        scale_minmax(r.tex, lo, hi);
        r.u = (uint8_t)rng.range(64, 192); r.v = (uint8_t)rng.range(64, 192);
        int x = rng.range(0, W - r.w), y = rng.range(0, H - r.h);
        int vx = rng.range(-V, V), vy = rng.range(-V, V);
        // drawn only when enabled, so sequences without zoom or rotation keep their random stream
        if (p->rect_zoom > 0) r.zoom = (2.0 * rng.uniform() - 1.0) * p->rect_zoom;
        if (p->rect_rotate_deg > 0) r.rot = (2.0 * rng.uniform() - 1.0) * p->rect_rotate_deg * M_PI / 180.0;
        r.px.resize(F); r.py.resize(F);
        for (int f = 0; f < F; f++) {
            r.px[f] = x; r.py[f] = y;
            x += vx; y += vy;
            if (x < 0) { x = -x; vx = -vx; }
            if (x > W - r.w) { x = 2 * (W - r.w) - x; vx = -vx; }
            if (y < 0) { y = -y; vy = -vy; }
            if (y > H - r.h) { y = 2 * (H - r.h) - y; vy = -vy; }
            x = std::max(0, std::min(W - r.w, x)); y = std::max(0, std::min(H - r.h, y));
        }
        s->rects.push_back(std::move(r));
    }
    *out = s;
    return 0;
}

static void render(const riv_synth* s, int f, uint8_t* rgb) {
    int W = (int)s->p.width, H = (int)s->p.height, cw = W / 2, ch = H / 2;
    std::vector<uint8_t> Y((size_t)W * H), U((size_t)cw * ch), V((size_t)cw * ch);
    Rng rng(s->p.seed * 0x9E3779B97F4A7C15ull + (uint64_t)f * 0xD1B54A32D192ED03ull + 77);
    double ox = s->gx[f] * 0.5 + s->cut_dx[f], oy = s->gy[f] * 0.5 + s->cut_dy[f];
    double sig = s->p.noise_sigma;
    std::vector<float> yf((size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) yf[(size_t)y * W + x] = sample_wrap(s->bg, s->nbg, x + ox, y + oy);
    std::vector<float> uf((size_t)cw * ch), vf((size_t)cw * ch);
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < cw; x++) {
            uf[(size_t)y * cw + x] = sample_wrap(s->cu, s->nch, x + ox * 0.5, y + oy * 0.5);
            vf[(size_t)y * cw + x] = sample_wrap(s->cv, s->nch, x + ox * 0.5, y + oy * 0.5);
        }
    for (const auto& r : s->rects) {
        int px = r.px[f], py = r.py[f];
        if (r.zoom != 0 || r.rot != 0) {
            // the rectangle scaled by (1+zoom)^f and turned by rot*f about its centre; a pixel
            // belongs to it when its inverse-mapped position lies inside the w x h texture window
            const double sc = std::pow(1.0 + r.zoom, f), a = r.rot * f, ca = std::cos(a), sa = std::sin(a);
            const double cx = px + 0.5 * r.w, cy = py + 0.5 * r.h, rad = 0.5 * sc * std::sqrt((double)r.w * r.w + (double)r.h * r.h);
            auto inv = [&](double x, double y, double* tx, double* ty) {
                const double u = x - cx, v = y - cy;
                *tx = (ca * u + sa * v) / sc + 0.5 * r.w;
                *ty = (-sa * u + ca * v) / sc + 0.5 * r.h;
                return *tx >= 0 && *tx < r.w && *ty >= 0 && *ty < r.h;
            };
            const int x0 = std::max(0, (int)std::floor(cx - rad)), x1 = std::min(W - 1, (int)std::ceil(cx + rad));
            const int y0 = std::max(0, (int)std::floor(cy - rad)), y1 = std::min(H - 1, (int)std::ceil(cy + rad));
            for (int y = y0; y <= y1; y++)
                for (int x = x0; x <= x1; x++) {
                    double tx, ty;
                    if (inv(x + 0.5, y + 0.5, &tx, &ty)) yf[(size_t)y * W + x] = sample_wrap(r.tex, s->nrt, tx, ty);
                }
            for (int y = y0 / 2; y <= y1 / 2 && y < ch; y++)
                for (int x = x0 / 2; x <= x1 / 2 && x < cw; x++) {
                    double tx, ty;
                    if (inv(2 * x + 1.0, 2 * y + 1.0, &tx, &ty)) { uf[(size_t)y * cw + x] = r.u; vf[(size_t)y * cw + x] = r.v; }
                }
            continue;
        }
        for (int y = py; y < py + r.h; y++)
            for (int x = px; x < px + r.w; x++) yf[(size_t)y * W + x] = r.tex[(size_t)((y - py) % s->nrt) * s->nrt + (x - px) % s->nrt];
        for (int y = (py + 1) / 2; y < (py + r.h) / 2; y++)
            for (int x = (px + 1) / 2; x < (px + r.w) / 2; x++) { uf[(size_t)y * cw + x] = r.u; vf[(size_t)y * cw + x] = r.v; }
    }
    for (size_t i = 0; i < Y.size(); i++) Y[i] = clip8(yf[i] + sig * rng.gauss());
    for (size_t i = 0; i < U.size(); i++) { U[i] = clip8(uf[i] + sig * rng.gauss()); V[i] = clip8(vf[i] + sig * rng.gauss()); }
    // BT.709 limited-range inverse, 14-bit fixed point, nearest chroma (same as the codec's)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int yy = Y[(size_t)y * W + x] - 16, uu = U[(size_t)(y >> 1) * cw + (x >> 1)] - 128, vv = V[(size_t)(y >> 1) * cw + (x >> 1)] - 128;
            int R = (19077 * yy + 29372 * vv + (1 << 13)) >> 14;
            int G = (19077 * yy - 3494 * uu - 8731 * vv + (1 << 13)) >> 14;
            int B = (19077 * yy + 34610 * uu + (1 << 13)) >> 14;
            uint8_t* p = rgb + ((size_t)y * W + x) * 3;
            p[0] = (uint8_t)std::min(255, std::max(0, R)); p[1] = (uint8_t)std::min(255, std::max(0, G)); p[2] = (uint8_t)std::min(255, std::max(0, B));
        }
}

int riv_synth_frames(const riv_synth* s, uint32_t start, uint32_t count, uint8_t* rgb_out, int threads) {
    if (start + count > s->p.frames) return 1;
    size_t fsz = (size_t)s->p.width * s->p.height * 3;
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([=]() { for (uint32_t i = t; i < count; i += threads) render(s, (int)(start + i), rgb_out + fsz * i); });
    for (auto& th : pool) th.join();
    return 0;
}

void riv_synth_destroy(riv_synth* s) { delete s; }

}  // extern "C"
