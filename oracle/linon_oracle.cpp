// linon CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A plain CPU restatement of linon's compute pass: the `main_view` entry of
// src/main_view.wgsl and the `overlay_desktop` entry of src/overlay.wgsl, templated on the
// scalar type so that one source gives
//   * the FP32 restatement (every WGSL operation rounded once to f32, evaluated in the
//     order the WGSL text writes it; built with -ffp-contract=off so that no FMA fuses two
//     roundings), used as the parity oracle of the HIP kernels, and
//   * the FP64 instantiation, used as the accuracy reference both FP32 results are
//     measured against.
// It is deliberately naive: brute force over all faces in index order, exactly as
// `ray_color` (src/main_view.wgsl:177-220) does it.  Nothing in the product library
// links or calls this file; only tests/, the smoke check and bench.py's CPU-baseline leg
// load it (through ctypes) as the checker.
//
// Parity pins for each function live in tests/test_oracle_*.py; the functions without a
// pin independent of this file are listed in the module docstring of oracle/oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------------
// vec3 arithmetic with WGSL semantics (component-wise, one rounding per operation)
// ---------------------------------------------------------------------------------------
template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> inline V3<T> operator-(V3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return mk(s * a.x, s * a.y, s * a.z); }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Rounding probe.  WGSL fixes exp/sin/cos/sqrt and division only to a few ulp, so two conforming
// FP32 evaluations of one ray may differ in their last bits.  With g_nudge = +1 or -1 the FP32
// instantiation moves every exp, sin, cos, sqrt and division result of the camera, field and
// step arithmetic one ulp up or down; a pixel whose end point then moves by more than the
// parity tolerance is rounding-sensitive.  The FP64 instantiation and g_nudge = 0 are untouched.
int g_nudge = 0;
template <typename T> inline T nudged(T r) { return r; }
template <> inline float nudged<float>(float r) {
    return g_nudge == 0 ? r : std::nextafter(r, g_nudge > 0 ? INFINITY : -INFINITY);
}
template <typename T> inline T t_sqrt(T x) { return nudged(T(std::sqrt(x))); }
template <typename T> inline T t_exp(T x) { return nudged(T(std::exp(x))); }
template <typename T> inline T t_sin(T x) { return nudged(T(std::sin(x))); }
template <typename T> inline T t_cos(T x) { return nudged(T(std::cos(x))); }
// division, which WGSL fixes only to 2.5 ulp, is probed the same way
template <typename T> inline T t_div(T a, T b) { return nudged(T(a / b)); }
template <typename T> inline V3<T> operator/(V3<T> a, T s) { return mk(t_div(a.x, s), t_div(a.y, s), t_div(a.z, s)); }
// the RK4 average's division by 6 (the kernels compute it correctly rounded in every build)
template <typename T> inline V3<T> div_exact(V3<T> a, T s) { return mk(a.x / s, a.y / s, a.z / s); }
template <typename T> inline T length(V3<T> a) { return t_sqrt(dot(a, a)); }
template <typename T> inline V3<T> normalize(V3<T> a) { return a / length(a); }
template <typename T> inline V3<T> vabs(V3<T> a) { return mk(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
// WGSL clamp(e, lo, hi) = min(max(e, lo), hi); the reference calls it as clamp(lo, hi, x).
template <typename T> inline T wgsl_clamp(T e, T lo, T hi) { return std::min(std::max(e, lo), hi); }
// WGSL reflect(e1, e2) = e1 - 2 * dot(e2, e1) * e2
template <typename T> inline V3<T> reflect(V3<T> e1, V3<T> e2) { return e1 - (T(2) * dot(e2, e1)) * e2; }

// ---------------------------------------------------------------------------------------
// Field-function helpers, src/main_view.wgsl:78-146
// ---------------------------------------------------------------------------------------
// mat3x3 (column-major) * v, written out column by column as the WGSL matrices are.
template <typename T> inline V3<T> matcols(V3<T> c0, V3<T> c1, V3<T> c2, V3<T> v) {
    return (c0 * v.x + c1 * v.y) + c2 * v.z;
}
template <typename T> V3<T> rotateX(V3<T> v, T phi) {  // main_view.wgsl:78-84
    T c = t_cos(phi), s = t_sin(phi);
    return matcols(mk<T>(1, 0, 0), mk<T>(0, c, s), mk<T>(0, -s, c), v);
}
template <typename T> V3<T> rotateY(V3<T> v, T phi) {  // main_view.wgsl:86-92
    T c = t_cos(phi), s = t_sin(phi);
    return matcols(mk<T>(c, 0, -s), mk<T>(0, 1, 0), mk<T>(s, 0, c), v);
}
template <typename T> V3<T> rotateZ(V3<T> v, T phi) {  // main_view.wgsl:94-100 (third row is zero)
    T c = t_cos(phi), s = t_sin(phi);
    return matcols(mk<T>(c, s, 0), mk<T>(-s, c, 0), mk<T>(0, 0, 0), v);
}
template <typename T> V3<T> translate(V3<T> v, T dx, T dy, T dz) {  // main_view.wgsl:102-108
    return mk(v.x + dx, v.y + dy, v.z + dz);
}
// Zhao et al. air refraction index, main_view.wgsl:111-117.  pow(10.0, -10.0) has constant
// arguments and is folded at shader creation, so it enters as the literal 1e-10.
template <typename T> T refraction_index(T t) {
    const T air_pressure = T(101325.0);
    const T c1 = T(0.0000104);
    const T c2 = T(0.00366);
    return t_div(c1 * air_pressure * (T(1) + air_pressure * (T(60.1) - T(0.972) * t) * T(1e-10)),
                 T(1) + c2 * t);
}
template <typename T> V3<T> refraction(T t_in, T t_out, V3<T> v_in, V3<T> n) {  // :119-136
    T eta_in = refraction_index(t_in);
    T eta_out = refraction_index(t_out);
    T cosi = wgsl_clamp(T(-1), T(1), dot(v_in, n));
    V3<T> n_ref = n;
    if (cosi < T(0)) {
        cosi = -cosi;
    } else {
        n_ref = -n;
    }
    T eta = t_div(eta_in, eta_out);
    T k = T(1) - eta * eta * (T(1) - cosi * cosi);
    if (k < T(0)) return reflect(v_in, n_ref);
    return eta * v_in + (eta * cosi - t_sqrt(k)) * n_ref;
}
template <typename T> T point_plane_distance(V3<T> p, V3<T> n, V3<T> p0) {  // :138-142
    T d = dot(p0, n);
    return std::fabs(dot(p, n) - d);
}
template <typename T> T sigmoid(T x) { return t_div(T(1), T(1) + t_exp(-x)); }  // :144-146

// ---------------------------------------------------------------------------------------
// Field bodies.  Ids 1..9 follow the enum order of src/functions.rs:2-13 (0 = Custom has no
// body).  16 is the stub body of main_view.wgsl:148.  17..20 are the four benchmark fields
// (swirl, lens, hot-air, warp), written in the same WGSL helper vocabulary.
// ---------------------------------------------------------------------------------------
template <typename T>
V3<T> mirage_temperatures(int id, V3<T> p_prev, V3<T> p, T* t_in, T* t_out, V3<T>* normal) {
    const T t_env = T(15);
    const bool spherical = (id == 1 || id == 3);
    const bool sig = (id == 3 || id == 4);
    const T t_src = spherical ? T(200) : T(30);
    const T max_dist = spherical ? T(0.25) : T(0.01);
    T dist_in, dist_out;
    if (spherical) {  // functions.rs:36-50, 67-81
        V3<T> center = mk<T>(T(-0.5), T(0.5), T(-0.5));
        V3<T> center_dest = p - center;
        *normal = normalize(center_dest);
        dist_in = length(p_prev - center);
        dist_out = length(center_dest);
    } else {  // functions.rs:52-65, 83-96
        V3<T> plane_p0 = mk<T>(T(0), T(0.1), T(0));
        V3<T> plane_n = mk<T>(T(0), T(1), T(0));
        *normal = plane_n;
        dist_in = point_plane_distance(p_prev, plane_n, plane_p0);
        dist_out = point_plane_distance(p, plane_n, plane_p0);
    }
    T part_in, part_out;
    if (sig) {
        part_in = wgsl_clamp(T(0), T(1), sigmoid(t_div(dist_in, max_dist) * T(12) - T(6)));
        part_out = wgsl_clamp(T(0), T(1), sigmoid(t_div(dist_out, max_dist) * T(12) - T(6)));
    } else {
        part_in = wgsl_clamp(T(0), T(1), t_div(dist_in, max_dist));
        part_out = wgsl_clamp(T(0), T(1), t_div(dist_out, max_dist));
    }
    *t_in = part_in * t_env + (T(1) - part_in) * t_src;
    *t_out = part_out * t_env + (T(1) - part_out) * t_src;
    return *normal;
}

template <typename T> T hot_air_temperature(V3<T> q) {
    V3<T> c = mk<T>(T(0), T(-0.5), T(0));
    return T(20) + T(80) * sigmoid(T(8) * (T(0.5) - length(q - c)));
}

template <typename T>
V3<T> field_eval(int id, V3<T> p_prev, V3<T> p, V3<T> v0, V3<T> v, T t) {
    const T PI = T(3.141592653589793);
    switch (id) {
        case 1: case 2: case 3: case 4: {
            T t_in, t_out;
            V3<T> n;
            mirage_temperatures(id, p_prev, p, &t_in, &t_out, &n);
            return refraction(t_in, t_out, v, n);
        }
        case 5: return translate(v0, T(0.5) * t, T(0), T(0));              // functions.rs:98-101
        case 6: return translate(v0, T(0), T(0), T(0.5) * t);              // functions.rs:103-106
        case 7: return normalize(v0 + rotateZ(v0, PI * t));                // functions.rs:108
        case 8: {                                                          // functions.rs:109-116
            const T rho = T(28), sigma = T(10), beta = T(8) / T(3);
            return mk(sigma * (p.y - p.x), p.x * (rho - p.z) - p.y, p.x * p.y - beta * p.z);
        }
        case 9: {                                                          // functions.rs:118-125
            const T a = T(0.1), b = T(0.1), c = T(14);
            return mk(-p.y - p.z, p.x + a * p.y, b + p.z * (p.x - c));
        }
        case 16: return v;                                                 // main_view.wgsl:148
        case 17: return rotateY(v0, T(0.5) * length(p));                   // swirl
        case 18: {                                                         // lens
            T r = length(p);
            return v - T(0.3) * p / std::max(r * r * r, T(0.001));
        }
        case 19: {                                                         // hot air
            V3<T> c = mk<T>(T(0), T(-0.5), T(0));
            return refraction(hot_air_temperature(p_prev), hot_air_temperature(p), v, normalize(p - c));
        }
        case 20:                                                           // warp
            return v0 + T(0.3) * mk(t_sin(T(4) * p.y), T(0), t_sin(T(4) * p.x));
        default: return v;
    }
}

bool field_known(int id) { return (id >= 1 && id <= 9) || (id >= 16 && id <= 20); }

// ---------------------------------------------------------------------------------------
// Scene + intersection, src/main_view.wgsl:150-220
// ---------------------------------------------------------------------------------------
template <typename T>
struct Scene {
    std::vector<V3<T>> v;
    std::vector<uint32_t> f;  // 3 per face
    uint32_t nf = 0;
    // per-face bounding boxes, widened by `pad`, for the optional exact prefilter of ray_color
    std::vector<T> box;  // 6 per face
    // the union of the boxes of each run of kPrefilterBlock consecutive faces (6 per block)
    std::vector<T> block_box;
    T pad = T(0);
};
constexpr uint32_t kPrefilterBlock = 256;

// Prefilter switch: skip the Moeller-Trumbore test of a face whose (slightly widened) bounding
// box misses the segment's bounding box, and a whole run of consecutive faces whose box union
// misses it.  Such a face cannot produce 0 < t < max_dist, and the faces left are still visited
// in ascending order, so the result is unchanged; it only saves time on large meshes.  tests/
// checks on == off.
bool g_prefilter = true;

template <typename T>
T hit_triangle(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> origin, V3<T> direction) {  // :150-175
    const T eps = T(0.0000001);
    V3<T> edge1 = v1 - v0;
    V3<T> edge2 = v2 - v0;
    V3<T> h = cross(direction, edge2);
    T a = dot(edge1, h);
    if (a > -eps && a < eps) return T(-1);  // backface_culling = false (:65)
    T f = t_div(T(1), a);
    V3<T> s = origin - v0;
    T u = f * dot(s, h);
    if (u < T(0) || u > T(1)) return T(-1);
    V3<T> q = cross(s, edge1);
    T vv = f * dot(direction, q);
    if (vv < T(0) || u + vv > T(1)) return T(-1);
    T t = f * dot(edge2, q);
    if (t > eps) return t;
    return T(-1);
}

template <typename T>
struct Hit {
    T color[4];
    int face;
};

// ray_color (:177-220): closest hit in [0, max_dist) over all faces in index order (strict <,
// so the lowest face index wins exact ties), two-sided camera-as-light Blinn-Phong.
template <typename T>
Hit<T> ray_color(const Scene<T>& sc, V3<T> origin, V3<T> direction, T max_dist, long long* tests) {
    Hit<T> r{{T(0), T(0), T(0), T(0)}, -1};
    T t = T(-1);
    V3<T> d1{}, d2{};
    int best = -1;
    const bool pre = g_prefilter && !sc.box.empty();
    V3<T> end = origin + direction * max_dist;
    const T slo[3] = {std::min(origin.x, end.x), std::min(origin.y, end.y), std::min(origin.z, end.z)};
    const T shi[3] = {std::max(origin.x, end.x), std::max(origin.y, end.y), std::max(origin.z, end.z)};
    long long tested = 0;
    for (uint32_t i = 0; i < sc.nf; ++i) {
        if (pre && i % kPrefilterBlock == 0) {
            const T* bb = &sc.block_box[6 * (size_t)(i / kPrefilterBlock)];
            if (bb[0] > shi[0] || bb[3] < slo[0] || bb[1] > shi[1] || bb[4] < slo[1] || bb[2] > shi[2] ||
                bb[5] < slo[2]) {
                i += kPrefilterBlock - 1;  // the loop's ++i moves to the next block
                continue;
            }
        }
        if (pre) {
            const T* bx = &sc.box[6 * (size_t)i];
            if (bx[0] > shi[0] || bx[3] < slo[0] || bx[1] > shi[1] || bx[4] < slo[1] || bx[2] > shi[2] ||
                bx[5] < slo[2])
                continue;
        }
        ++tested;
        V3<T> a = sc.v[sc.f[3 * i]], b = sc.v[sc.f[3 * i + 1]], c = sc.v[sc.f[3 * i + 2]];
        T t_new = hit_triangle(a, b, c, origin, direction);
        if (t_new > T(0) && t_new < max_dist && (t < T(0) || t_new < t)) {
            t = t_new;
            d1 = b - a;
            d2 = c - a;
            best = (int)i;
        }
    }
    if (tests) *tests += pre ? tested : (long long)sc.nf;
    if (t > T(0) && t < max_dist) {
        V3<T> normal = normalize(cross(d1, d2));
        T intensity = std::max(dot(normal, -direction), T(0));
        if (intensity == T(0)) intensity = std::max(dot(-normal, -direction), T(0));
        T spec = std::pow(intensity, T(64));
        V3<T> an = vabs(normal);
        // (ambient + diffuse + specular) * abs(normal), light colour (1,1,1), ambient 0.01
        T s = (T(0.01) + intensity) + spec;
        r.color[0] = s * an.x;
        r.color[1] = s * an.y;
        r.color[2] = s * an.z;
        r.color[3] = t;
        r.face = best;
    }
    return r;
}

struct CameraIn {
    float origin[4];
    float view_direction[4];  // .w: vertical field of view in degrees, 0 = fixed 0.025 viewport
    float up[4];
    float view_matrix[16];
};
struct SettingsIn {
    float field_weight;
    float mouse_pos[2];
    int32_t overlay_mode;
    int32_t central_difference_delta;
    float lyapunov_scaling;
};
struct IntegratorIn {
    float h_initial;
    float t_max;
    int32_t adaptive;
    int32_t max_steps;
    int32_t mode;  // 0 = reference loop (t <= t_max), 1 = escape (leave sphere / first hit)
    float escape_radius;
    int32_t samples_per_ray;
    int32_t sample_steps;
    int32_t need_mapping;  // escape mode only: keep integrating after a hit (overlay wants the end point)
    int32_t outline_rays;  // layout only: outline positions are passed to sample_rays explicitly
    int32_t clip_exit;     // escape mode: a ray leaving the sphere from inside ends on it
    int32_t pad;
};

template <typename T>
struct CameraBasis {
    V3<T> origin, view_direction, nh, nv;
    T focal, vh, vw;
};

// main_view.wgsl:347-366
template <typename T>
CameraBasis<T> camera_basis(const CameraIn& c, int W, int H) {
    CameraBasis<T> b;
    b.origin = mk<T>(c.origin[0], c.origin[1], c.origin[2]);
    b.view_direction = mk<T>(c.view_direction[0], c.view_direction[1], c.view_direction[2]);
    V3<T> up = mk<T>(c.up[0], c.up[1], c.up[2]);
    T width = T(W), height = T(H);
    T aspect = width / height;
    b.focal = T(0.035);
    b.vh = T(0.025);
    if (c.view_direction[3] > 0.0f) {
        // extension: a pinhole with the given vertical FOV at the same focal length
        b.vh = T(2) * b.focal * T(std::tan(double(c.view_direction[3]) * 3.141592653589793 / 360.0));
    }
    b.vw = aspect * b.vh;
    V3<T> w = -b.view_direction;
    V3<T> horizontal = cross(up, w);
    V3<T> vertical = cross(w, horizontal);
    b.nh = normalize(horizontal);
    b.nv = normalize(vertical);
    return b;
}

template <typename T>
V3<T> camera_ray(const CameraBasis<T>& b, int W, int H, int gx, int gy) {
    T width = T(W), height = T(H);
    T u = t_div(T(gx), width - T(1)) * b.vw - T(0.5) * b.vw;
    T v = t_div(T(gy), height - T(1)) * b.vh - T(0.5) * b.vh;
    V3<T> s = u * b.nh + v * b.nv + b.focal * b.view_direction;
    return normalize(s);
}

template <typename T>
struct RayOut {
    T color[4];
    T mapping[4];
    int face;
    int steps;   // RK4 iterations evaluated (accepted + adaptive redos)
    long long tests;
    int flags;   // bit 0: an adaptive-step comparison was decided at rounding-noise level
    T end_dir[3];  // cur_dir at the end of the loop
};

// Escape mode with clip_exit: the crossing of the last step [p0, p0 + s] with the sphere
// |x|^2 = r2 when the step started inside it, as the larger root u in [0, 1] of
// |s|^2 u^2 + 2 (p0 . s) u + (|p0|^2 - r2) = 0 (no counterpart in the reference).
template <typename T> V3<T> sphere_exit(V3<T> p0, V3<T> s, T r2) {
    T c = dot(p0, p0) - r2;
    if (!(c <= T(0))) return p0 + s;
    T a = dot(s, s);
    T b = dot(p0, s);
    T root = t_sqrt(b * b - a * c);
    T u = b > T(0) ? t_div(-c, b + root) : t_div(root - b, a);
    u = std::fmin(std::fmax(u, T(0)), T(1));
    return p0 + u * s;
}

// nonlinear_ray_color, main_view.wgsl:229-279 (mode 0), plus the escape termination.
template <typename T>
RayOut<T> trace(const Scene<T>& sc, int field, const IntegratorIn& ig, T field_weight,
                V3<T> start_point, V3<T> start_dir) {
    RayOut<T> out{};
    bool has_color = false;
    Hit<T> col{{T(0), T(0), T(0), T(0)}, -1};
    V3<T> cur_point = start_point, cur_dir = start_dir;
    T t = T(0);
    V3<T> last_v = mk<T>(0, 0, 0);
    T last_diff = T(-1);
    const T h_initial = T(ig.h_initial);
    T h = h_initial;
    const T r2 = T(ig.escape_radius) * T(ig.escape_radius);
    int iters = 0;
    long long tests = 0;
    const T half = T(0.5);
    // The loop bound is the reference's f32 running sum `t = t + h` (main_view.wgsl:244,265);
    // the FP64 instantiation keeps that f32 counter for the bound so both precisions take the
    // same number of steps and their difference is integration error only.
    float t_loop = 0.0f;
    while (iters < ig.max_steps) {
        if (ig.mode == 0 && !(t_loop <= ig.t_max)) break;
        ++iters;
        V3<T> k1 = field_eval(field, cur_point, cur_point, start_dir, cur_dir, t);
        V3<T> k2 = field_eval(field, cur_point, cur_point + (half * h) * k1, start_dir, k1, t + half * h);
        V3<T> k3 = field_eval(field, cur_point, cur_point + (half * h) * k2, start_dir, k2, t + half * h);
        V3<T> k4 = field_eval(field, cur_point, cur_point + h * k3, start_dir, k3, t + h);
        V3<T> v = div_exact(((k1 + T(2) * k2) + T(2) * k3) + k4, T(6));
        T diff = length(v - last_v);
        if (ig.adaptive) {
            // Conditioning of the adaptive rule (main_view.wgsl:252,267): where the field is
            // uniform, V - last_v is rounding noise and the comparison's outcome depends on the
            // last bits of the arithmetic; flag such rays (and near-ties) for the parity tests.
            const T noise = T(1e-5) * std::max(length(v), T(1e-30));
            const bool trig_test = last_diff >= T(0) && h > T(0.002);
            const bool rest_test = h < h_initial && last_diff >= T(0);
            if ((trig_test || rest_test) && (diff < noise || last_diff < noise)) out.flags |= 1;
            if (trig_test && std::fabs(diff - T(10) * last_diff) <= T(1e-4) * diff) out.flags |= 1;
            if (rest_test && std::fabs(T(2) * diff - last_diff) <= T(1e-4) * last_diff) out.flags |= 1;
        }
        if (ig.adaptive && last_diff >= T(0) && diff > T(10) * last_diff && h > T(0.002)) {
            h = T(0.001);
            continue;
        }
        cur_dir = (T(1) - field_weight) * cur_dir + field_weight * v;
        V3<T> step_dir = cur_dir * h;
        if (!has_color) {
            col = ray_color(sc, cur_point, normalize(step_dir), length(step_dir), &tests);
            has_color = col.color[3] > T(0);
        }
        const V3<T> prev_point = cur_point;
        cur_point = cur_point + step_dir;
        t = t + h;
        t_loop = t_loop + (float)h;
        last_v = v;
        if (T(2) * diff < last_diff && h < h_initial) h = h_initial;
        last_diff = diff;
        if (ig.mode == 1) {
            if (has_color && !ig.need_mapping) break;
            if (dot(cur_point, cur_point) > r2 && dot(cur_point, cur_dir) > T(0)) {
                if (ig.clip_exit) cur_point = sphere_exit(prev_point, step_dir, r2);
                break;
            }
        }
    }
    out.mapping[0] = cur_point.x;
    out.mapping[1] = cur_point.y;
    out.mapping[2] = cur_point.z;
    out.mapping[3] = T(1);
    out.end_dir[0] = cur_dir.x;
    out.end_dir[1] = cur_dir.y;
    out.end_dir[2] = cur_dir.z;
    if (has_color) {
        for (int i = 0; i < 3; ++i) out.color[i] = col.color[i];
        out.face = col.face;
    } else {
        out.color[0] = out.color[1] = out.color[2] = T(0);
        out.face = -1;
    }
    out.color[3] = T(1);
    out.steps = iters;
    out.tests = tests;
    return out;
}

inline uint8_t unorm8(double x) {
    // Vulkan float -> UNORM8: clamp to [0,1], scale by 255, round to nearest
    if (!(x > 0.0)) return 0;  // also maps NaN to 0
    if (x >= 1.0) return 255;
    return (uint8_t)std::floor(x * 255.0 + 0.5);
}

template <typename T>
Scene<T> make_scene(const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf) {
    Scene<T> sc;
    sc.v.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) sc.v[i] = mk<T>(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    sc.f.assign(faces, faces + 3 * (size_t)nf);
    sc.nf = nf;
    T ext = T(0);
    sc.box.resize(6 * (size_t)nf);
    for (uint32_t i = 0; i < nf; ++i) {
        T* bx = &sc.box[6 * (size_t)i];
        for (int k = 0; k < 3; ++k) {
            T a = verts[3 * (size_t)faces[3 * i] + k], b = verts[3 * (size_t)faces[3 * i + 1] + k],
              c = verts[3 * (size_t)faces[3 * i + 2] + k];
            bx[k] = std::min(a, std::min(b, c));
            bx[3 + k] = std::max(a, std::max(b, c));
            ext = std::max(ext, std::max(std::fabs(bx[k]), std::fabs(bx[3 + k])));
        }
    }
    sc.pad = T(1e-4) * (ext + T(1));
    for (uint32_t i = 0; i < nf; ++i)
        for (int k = 0; k < 3; ++k) {
            sc.box[6 * (size_t)i + k] -= sc.pad;
            sc.box[6 * (size_t)i + 3 + k] += sc.pad;
        }
    sc.block_box.assign(6 * (size_t)((nf + kPrefilterBlock - 1) / kPrefilterBlock), T(0));
    for (uint32_t i = 0; i < nf; ++i) {
        T* bb = &sc.block_box[6 * (size_t)(i / kPrefilterBlock)];
        const T* bx = &sc.box[6 * (size_t)i];
        for (int k = 0; k < 3; ++k) {
            bb[k] = i % kPrefilterBlock == 0 ? bx[k] : std::min(bb[k], bx[k]);
            bb[3 + k] = i % kPrefilterBlock == 0 ? bx[3 + k] : std::max(bb[3 + k], bx[3 + k]);
        }
    }
    return sc;
}

template <typename T>
int render_rows(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf, int W, int H,
                int row0, int row1, int nthreads, uint8_t* color8, T* color, T* mapping, int32_t* hit,
                int32_t* steps, long long* counters, uint8_t* flags) {
    if (!cam || !st || !ig || W < 2 || H < 2 || row0 < 0 || row1 > H || row0 > row1) return -1;
    if (!field_known(field)) return -2;
    for (uint32_t i = 0; i < 3 * (size_t)nf; ++i)
        if (faces[i] >= nv) return -3;
    Scene<T> sc = make_scene<T>(verts, nv, faces, nf);
    CameraBasis<T> cb = camera_basis<T>(*cam, W, H);
    const T w = T(st->field_weight);
    long long tot_steps = 0, tot_tests = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot_steps, tot_tests)
#endif
    for (int gy = row0; gy < row1; ++gy) {
        for (int gx = 0; gx < W; ++gx) {
            V3<T> dir = camera_ray(cb, W, H, gx, gy);
            RayOut<T> r = trace(sc, field, *ig, w, cb.origin, dir);
            size_t idx = (size_t)(H - 1 - gy) * W + gx;  // texel row flip, main_view.wgsl:342
            if (color8)
                for (int c = 0; c < 4; ++c) color8[4 * idx + c] = unorm8(double(r.color[c]));
            if (color)
                for (int c = 0; c < 4; ++c) color[4 * idx + c] = r.color[c];
            if (mapping)
                for (int c = 0; c < 4; ++c) mapping[4 * idx + c] = r.mapping[c];
            if (hit) hit[idx] = r.face;
            if (steps) steps[idx] = r.steps;
            if (flags) flags[idx] = (uint8_t)r.flags;
            tot_steps += r.steps;
            tot_tests += r.tests;
        }
    }
    if (counters) {
        counters[0] = tot_steps;
        counters[1] = tot_tests;
    }
    return 0;
}

// overlay.wgsl:4-40,99-112: sigma_max of the 3x2 central-difference matrix via the 2x2
// eigenvalues of G^T G, with the complex-root branch of vec2squareroots kept.
template <typename T>
T lyapunov(V3<T> x_next, V3<T> x_prev, V3<T> y_next, V3<T> y_prev) {
    V3<T> gx = T(0.5) * (x_next - x_prev);
    V3<T> gy = T(0.5) * (y_next - y_prev);
    // m = transpose(G) * G; column j = (dot(gx, G_j), dot(gy, G_j))
    T m00 = dot(gx, gx), m01 = dot(gy, gx);  // column 0
    T m10 = dot(gx, gy), m11 = dot(gy, gy);  // column 1
    T det = m00 * m11 - m01 * m10;           // mat2det
    T a0 = det, a1 = -(m00 + m11);           // mat2invariants
    T e0, e1;
    T discrim = a1 * a1 - T(4) * a0;         // vec2squareroots
    if (discrim >= T(0)) {
        T root = std::sqrt(discrim);
        e0 = (-a1 - root) / T(2);
        e1 = (-a1 + root) / T(2);
    } else {
        T root = std::sqrt(-discrim);
        e0 = -a1 / T(2);
        e1 = root / T(2);
    }
    return std::sqrt(std::max(e0, e1));
}

// srgb_from_linear, overlay.wgsl:148-154: per channel, x / 12.92 below 0.0031308, else
// 1.055 x^(1/2.4) - 0.055 (1.0 / 2.4 is an f32 constant expression in WGSL).
template <typename T>
T srgb_from_linear(T x) {
    const T inv = T(1.0f / 2.4f);
    return x < T(0.0031308f) ? x / T(12.92f) : T(1.055f) * std::pow(x, inv) - T(0.055f);
}

// overlay_desktop (overlay.wgsl:137-145) and, with web, overlay_web (:156-165), which stores the
// sRGB encoding of the colour's rgb (alpha unchanged); the reference picks overlay_web for its
// wasm32 build (main_view.rs:422-426).
template <typename T>
int overlay(const SettingsIn* st, int W, int H, const uint8_t* ray_casting, const T* mapping,
            uint8_t* final8, float* exponents, float* lyap, bool web = false) {
    if (!st || W < 1 || H < 1) return -1;
    const int delta = st->central_difference_delta;
    const int padding = delta * 2;
    for (int cy = 0; cy < H; ++cy) {
        for (int cx = 0; cx < W; ++cx) {
            size_t idx = (size_t)cy * W + cx;
            T color[4];
            for (int c = 0; c < 4; ++c) color[c] = T(ray_casting[4 * idx + c]) / T(255);
            if (lyap) lyap[idx] = -1.0f;
            T out[4] = {color[0], color[1], color[2], color[3]};
            bool inside = !(cx < padding || cy < padding || cx >= W - padding || cy >= H - padding);
            if (st->overlay_mode != 0 && inside && st->overlay_mode == 1) {
                auto m = [&](int x, int y) {
                    const T* p = mapping + 4 * ((size_t)y * W + x);
                    return mk<T>(p[0], p[1], p[2]);
                };
                T lam = lyapunov(m(cx + delta, cy), m(cx - delta, cy), m(cx, cy + delta), m(cx, cy - delta));
                T s = std::exp(T(st->lyapunov_scaling) * lam - T(5));
                size_t eidx = (size_t)H * cy + cx;  // row stride = height, overlay.wgsl:128
                if (exponents && eidx < (size_t)W * H) exponents[eidx] = float(s);
                if (lyap) lyap[idx] = float(lam);
                if (s >= T(0.8)) {
                    T alpha = std::min(T(1), s) * T(0.5);
                    for (int c = 0; c < 4; ++c) out[c] = alpha * T(1) + (T(1) - alpha) * color[c];
                }
            }
            if (web)
                for (int c = 0; c < 3; ++c) out[c] = srgb_from_linear(out[c]);
            for (int c = 0; c < 4; ++c) final8[4 * idx + c] = unorm8(double(out[c]));
        }
    }
    return 0;
}

// sample_rays, main_view.wgsl:281-314 with the launch-side ray selection of :470-479.
const float kSampleOffsets[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};
const float kSampleColors[8][3] = {{1, 0, 0},     {0.5f, 0.5f, 0}, {0, 1, 0}, {0.5f, 1, 0},
                                   {1, 1, 0},     {0.5f, 0.5f, 0.5f}, {0, 0, 1}, {0.5f, 0, 0.5f}};

// The outline-ray searches of main_view.wgsl:381-468, loop for loop: 8 scans of the exponents
// buffer (index y * H + x, the overlay's row stride H) for entries >= 0.8.  An index at or past
// W * H (search 1's rows y >= H, or x >= H aliasing past the end when H > W) reads as 0 here:
// WGSL leaves out-of-bounds reads of a storage buffer to the implementation.  out: 8 x {x, y,
// found}; not found leaves pos = (0, 0) as the WGSL does.
void outline_search(const float* ex, int W, int H, int32_t* out) {
    const float min_exp = 0.8f;
    auto e = [&](int x, int y) -> float {
        const long long idx = (long long)y * H + x;
        return (idx >= 0 && idx < (long long)W * H) ? ex[idx] : 0.0f;
    };
    for (int r = 0; r < 8; ++r) {
        int px = 0, py = 0, sum = 0;
        bool found = false, any = false;
        if (r == 0 || r == 2 || r == 4 || r == 6) {  // corners: key >= sum keeps the last maximum
            for (int x = 0; x < W; ++x)
                for (int y = 0; y < H; ++y) {
                    int key = r == 0 ? (W - x) + y : r == 2 ? x + y : r == 4 ? x + (H - y) : (W - x) + (H - y);
                    if (e(x, y) >= min_exp && key >= sum) {
                        px = x;
                        py = y;
                        sum = key;
                        any = true;
                    }
                }
        } else if (r == 1) {  // bottom middle: y from W - 1 (sic) down, x up, first hit
            for (int y = W - 1; y >= 0 && !found; --y)
                for (int x = 0; x < W && !found; ++x)
                    if (e(x, y) >= min_exp) { px = x; py = y; found = true; }
        } else if (r == 3) {  // middle right: x from W - 1 down, y up
            for (int x = W - 1; x >= 0 && !found; --x)
                for (int y = 0; y < H && !found; ++y)
                    if (e(x, y) >= min_exp) { px = x; py = y; found = true; }
        } else if (r == 5) {  // top middle: y up, x up
            for (int y = 0; y < H && !found; ++y)
                for (int x = 0; x < W && !found; ++x)
                    if (e(x, y) >= min_exp) { px = x; py = y; found = true; }
        } else {  // middle left: x up, y up
            for (int x = 0; x < W && !found; ++x)
                for (int y = 0; y < H && !found; ++y)
                    if (e(x, y) >= min_exp) { px = x; py = y; found = true; }
        }
        out[3 * r] = px;
        out[3 * r + 1] = py;
        out[3 * r + 2] = (found || any) ? 1 : 0;
    }
}

template <typename T>
int sample_rays(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field, int W, int H,
                T* out, const int32_t* outline = nullptr) {
    if (!field_known(field)) return -2;
    CameraBasis<T> cb = camera_basis<T>(*cam, W, H);
    const T w = T(st->field_weight);
    const int steps = ig->sample_steps;
    const int nsamp = ig->samples_per_ray;
    const int stride = steps / nsamp;
    const T h = T(0.001);
    for (int r = 0; r < 8; ++r) {
        T px = T(st->mouse_pos[0]) + T(0.01) * T(kSampleOffsets[r][0]);
        T py = T(st->mouse_pos[1]) + T(0.01) * T(kSampleOffsets[r][1]);
        if (outline) {  // main_view.wgsl:466-467
            const T width = T(W), height = T(H);
            px = t_div(T(outline[3 * r]), width - T(1));
            py = t_div(height - T(1) - T(outline[3 * r + 1]), height - T(1));
        }
        T u2 = px * cb.vw - T(0.5) * cb.vw;
        T v2 = py * cb.vh - T(0.5) * cb.vh;
        V3<T> s2 = u2 * cb.nh + v2 * cb.nv + cb.focal * cb.view_direction;
        V3<T> start_dir = normalize(s2);
        V3<T> cur_point = cb.origin, cur_dir = start_dir;
        T t = T(0);
        for (int i = 0; i < steps; ++i) {
            V3<T> k1 = field_eval(field, cur_point, cur_point, start_dir, cur_dir, t);
            V3<T> k2 = field_eval(field, cur_point, cur_point + (T(0.5) * h) * k1, start_dir, k1, t + T(0.5) * h);
            V3<T> k3 = field_eval(field, cur_point, cur_point + (T(0.5) * h) * k2, start_dir, k2, t + T(0.5) * h);
            V3<T> k4 = field_eval(field, cur_point, cur_point + h * k3, start_dir, k3, t + h);
            V3<T> v = div_exact(((k1 + T(2) * k2) + T(2) * k3) + k4, T(6));
            cur_dir = (T(1) - w) * cur_dir + w * v;
            V3<T> step_dir = cur_dir * h;
            if (i % stride == 0 && i / stride < nsamp) {
                T* o = out + 8 * ((size_t)r * nsamp + i / stride);
                o[0] = cur_point.x; o[1] = cur_point.y; o[2] = cur_point.z; o[3] = T(1);
                o[4] = T(kSampleColors[r][0]); o[5] = T(kSampleColors[r][1]);
                o[6] = T(kSampleColors[r][2]); o[7] = T(0.5);
            }
            cur_point = cur_point + step_dir;
            t = t + h;
        }
    }
    return 0;
}

// Rays of an explicit pixel list (invocation coordinates gx, gy), for sampled checks and timings.
template <typename T>
int render_pixels(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field, const float* verts,
                  uint32_t nv, const uint32_t* faces, uint32_t nf, int W, int H, const int32_t* pix, int n,
                  int nthreads, T* mapping, int32_t* hit, int32_t* steps, long long* counters) {
    if (!cam || !st || !ig || !pix || W < 2 || H < 2 || n < 0) return -1;
    if (!field_known(field)) return -2;
    for (uint32_t i = 0; i < 3 * (size_t)nf; ++i)
        if (faces[i] >= nv) return -3;
    Scene<T> sc = make_scene<T>(verts, nv, faces, nf);
    CameraBasis<T> cb = camera_basis<T>(*cam, W, H);
    const T w = T(st->field_weight);
    long long tot_steps = 0, tot_tests = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot_steps, tot_tests)
#endif
    for (int i = 0; i < n; ++i) {
        V3<T> dir = camera_ray(cb, W, H, pix[2 * i], pix[2 * i + 1]);
        RayOut<T> r = trace(sc, field, *ig, w, cb.origin, dir);
        if (mapping)
            for (int c = 0; c < 4; ++c) mapping[4 * (size_t)i + c] = r.mapping[c];
        if (hit) hit[i] = r.face;
        if (steps) steps[i] = r.steps;
        tot_steps += r.steps;
        tot_tests += r.tests;
    }
    if (counters) {
        counters[0] = tot_steps;
        counters[1] = tot_tests;
    }
    return 0;
}

// Flow map of the field through the bounding sphere: from an entry state (p, q) — q is the
// initial direction v0 for fields that read v0, the current direction for those that read v —
// integrate in escape mode without a mesh until the ray leaves the sphere (or max_steps).
template <typename T>
void flowmap(int field, const IntegratorIn& ig, T field_weight, const T* in, T* out, int32_t* steps, int n,
             int nthreads) {
    Scene<T> empty;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int i = 0; i < n; ++i) {
        const T* x = in + 6 * (size_t)i;
        RayOut<T> r = trace(empty, field, ig, field_weight, mk(x[0], x[1], x[2]), mk(x[3], x[4], x[5]));
        T* o = out + 6 * (size_t)i;
        o[0] = r.mapping[0];
        o[1] = r.mapping[1];
        o[2] = r.mapping[2];
        o[3] = r.end_dir[0];
        o[4] = r.end_dir[1];
        o[5] = r.end_dir[2];
        if (steps) steps[i] = r.steps;
    }
}

}  // namespace

extern "C" {

int oracle_abi_version() { return 7; }
void oracle_set_nudge(int ulps) { g_nudge = ulps > 0 ? 1 : (ulps < 0 ? -1 : 0); }
void oracle_set_prefilter(int on) { g_prefilter = on != 0; }

int oracle_render_f32(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                      const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf, int W, int H,
                      int row0, int row1, int nthreads, uint8_t* color8, float* color, float* mapping,
                      int32_t* hit, int32_t* steps, long long* counters, uint8_t* flags) {
    return render_rows<float>(cam, st, ig, field, verts, nv, faces, nf, W, H, row0, row1, nthreads, color8,
                              color, mapping, hit, steps, counters, flags);
}
int oracle_render_f64(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                      const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf, int W, int H,
                      int row0, int row1, int nthreads, uint8_t* color8, double* color, double* mapping,
                      int32_t* hit, int32_t* steps, long long* counters, uint8_t* flags) {
    return render_rows<double>(cam, st, ig, field, verts, nv, faces, nf, W, H, row0, row1, nthreads, color8,
                               color, mapping, hit, steps, counters, flags);
}
int oracle_render_pixels_f32(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                             const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf, int W, int H,
                             const int32_t* pix, int n, int nthreads, float* mapping, int32_t* hit, int32_t* steps,
                             long long* counters) {
    return render_pixels<float>(cam, st, ig, field, verts, nv, faces, nf, W, H, pix, n, nthreads, mapping, hit,
                                steps, counters);
}
int oracle_render_pixels_f64(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                             const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf, int W, int H,
                             const int32_t* pix, int n, int nthreads, double* mapping, int32_t* hit, int32_t* steps,
                             long long* counters) {
    return render_pixels<double>(cam, st, ig, field, verts, nv, faces, nf, W, H, pix, n, nthreads, mapping, hit,
                                 steps, counters);
}
int oracle_overlay_f32(const SettingsIn* st, int W, int H, const uint8_t* ray_casting, const float* mapping,
                       uint8_t* final8, float* exponents, float* lyap) {
    return overlay<float>(st, W, H, ray_casting, mapping, final8, exponents, lyap);
}
int oracle_overlay_f64(const SettingsIn* st, int W, int H, const uint8_t* ray_casting, const double* mapping,
                       uint8_t* final8, float* exponents, float* lyap) {
    return overlay<double>(st, W, H, ray_casting, mapping, final8, exponents, lyap);
}
int oracle_overlay_web_f32(const SettingsIn* st, int W, int H, const uint8_t* ray_casting, const float* mapping,
                           uint8_t* final8, float* exponents, float* lyap) {
    return overlay<float>(st, W, H, ray_casting, mapping, final8, exponents, lyap, true);
}
float oracle_srgb_from_linear_f32(float x) { return srgb_from_linear<float>(x); }
int oracle_sample_rays_f32(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field, int W,
                           int H, float* out) {
    return sample_rays<float>(cam, st, ig, field, W, H, out);
}
int oracle_sample_rays_f64(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field, int W,
                           int H, double* out) {
    return sample_rays<double>(cam, st, ig, field, W, H, out);
}
void oracle_outline_search(const float* exponents, int W, int H, int32_t* out) {
    outline_search(exponents, W, H, out);
}
int oracle_sample_rays_outline_f32(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                                   int W, int H, const int32_t* outline, float* out) {
    return sample_rays<float>(cam, st, ig, field, W, H, out, outline);
}
int oracle_sample_rays_outline_f64(const CameraIn* cam, const SettingsIn* st, const IntegratorIn* ig, int field,
                                   int W, int H, const int32_t* outline, double* out) {
    return sample_rays<double>(cam, st, ig, field, W, H, out, outline);
}

// ---- single-function entry points used by the pinning tests --------------------------
float oracle_hit_triangle_f32(const float* tri9, const float* o, const float* d) {
    return hit_triangle<float>(mk(tri9[0], tri9[1], tri9[2]), mk(tri9[3], tri9[4], tri9[5]),
                               mk(tri9[6], tri9[7], tri9[8]), mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]));
}
double oracle_hit_triangle_f64(const double* tri9, const double* o, const double* d) {
    return hit_triangle<double>(mk(tri9[0], tri9[1], tri9[2]), mk(tri9[3], tri9[4], tri9[5]),
                                mk(tri9[6], tri9[7], tri9[8]), mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]));
}
// closest hit over a mesh: returns the face index (-1: none), writes t and the colour
int oracle_ray_color_f32(const float* verts, uint32_t nv, const uint32_t* faces, uint32_t nf, const float* o,
                         const float* d, float max_dist, float* rgba) {
    Scene<float> sc = make_scene<float>(verts, nv, faces, nf);
    Hit<float> h = ray_color(sc, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), max_dist, nullptr);
    for (int c = 0; c < 4; ++c) rgba[c] = h.color[c];
    return h.face;
}
int oracle_field_f64(int id, const double* p_prev, const double* p, const double* v0, const double* v, double t,
                     double* out) {
    if (!field_known(id)) return -2;
    V3<double> r = field_eval<double>(id, mk(p_prev[0], p_prev[1], p_prev[2]), mk(p[0], p[1], p[2]),
                                      mk(v0[0], v0[1], v0[2]), mk(v[0], v[1], v[2]), t);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 0;
}
int oracle_field_f32(int id, const float* p_prev, const float* p, const float* v0, const float* v, float t,
                     float* out) {
    if (!field_known(id)) return -2;
    V3<float> r = field_eval<float>(id, mk(p_prev[0], p_prev[1], p_prev[2]), mk(p[0], p[1], p[2]),
                                    mk(v0[0], v0[1], v0[2]), mk(v[0], v[1], v[2]), t);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 0;
}
double oracle_refraction_index_f64(double t) { return refraction_index<double>(t); }
int oracle_refraction_f64(double t_in, double t_out, const double* v, const double* n, double* out) {
    V3<double> r = refraction<double>(t_in, t_out, mk(v[0], v[1], v[2]), mk(n[0], n[1], n[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 0;
}
double oracle_lyapunov_f64(const double* xn, const double* xp, const double* yn, const double* yp) {
    return lyapunov<double>(mk(xn[0], xn[1], xn[2]), mk(xp[0], xp[1], xp[2]), mk(yn[0], yn[1], yn[2]),
                            mk(yp[0], yp[1], yp[2]));
}
int oracle_camera_ray_f64(const CameraIn* cam, int W, int H, int gx, int gy, double* out) {
    CameraBasis<double> cb = camera_basis<double>(*cam, W, H);
    V3<double> d = camera_ray(cb, W, H, gx, gy);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
    return 0;
}
// one ray through the integrator: returns the RK4 iteration count, writes the end point
int oracle_trace_f64(int field, const IntegratorIn* ig, double field_weight, const float* verts, uint32_t nv,
                     const uint32_t* faces, uint32_t nf, const double* o, const double* d, double* end_point,
                     int32_t* face) {
    Scene<double> sc = make_scene<double>(verts, nv, faces, nf);
    RayOut<double> r = trace(sc, field, *ig, field_weight, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]));
    for (int c = 0; c < 3; ++c) end_point[c] = r.mapping[c];
    if (face) *face = r.face;
    return r.steps;
}
int oracle_trace_f32(int field, const IntegratorIn* ig, float field_weight, const float* verts, uint32_t nv,
                     const uint32_t* faces, uint32_t nf, const float* o, const float* d, float* end_point,
                     int32_t* face) {
    Scene<float> sc = make_scene<float>(verts, nv, faces, nf);
    RayOut<float> r = trace(sc, field, *ig, field_weight, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]));
    for (int c = 0; c < 3; ++c) end_point[c] = r.mapping[c];
    if (face) *face = r.face;
    return r.steps;
}
int oracle_flowmap_f32(int field, const IntegratorIn* ig, float field_weight, const float* in, float* out,
                       int32_t* steps, int n, int nthreads) {
    if (!field_known(field) || !ig || n < 0) return -1;
    flowmap<float>(field, *ig, field_weight, in, out, steps, n, nthreads);
    return 0;
}
int oracle_flowmap_f64(int field, const IntegratorIn* ig, double field_weight, const double* in, double* out,
                       int32_t* steps, int n, int nthreads) {
    if (!field_known(field) || !ig || n < 0) return -1;
    flowmap<double>(field, *ig, field_weight, in, out, steps, n, nthreads);
    return 0;
}
int oracle_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
