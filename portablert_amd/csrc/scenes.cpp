// Deterministic synthetic workloads for the nearest-hit path.
//
// Every scene and ray set is a pure function of (workload, seed, index): scenes are built from
// one SplitMix64 stream per scene, rays from a counter-based stream keyed by the ray index, so
// any sub-range of a ray set can be regenerated on its own (the parity tests take samples of
// the 8M-ray sets this way). Arithmetic is done in double and rounded once to float; the build
// uses -ffp-contract=off so no fused operation changes a rounding between machines.
//
// Workloads (triangle layout is the reference's Tri, 9 floats v0 v1 v2, core.hpp:24-25; ray
// layout is the reference's Ray, origin then direction, core.hpp:19-22):
//   0  soup:      16,384 triangles, centroids U[-1,1]^3, vertex offsets U[-0.05,0.05]^3;
//                 1024x1024 pinhole from (0,0,-3) at the origin, 60 deg vertical FOV.
//   1  sphere:    icosphere subdivided 6 times (81,920 triangles), r=1, radial displacement
//                 0.08 * value-noise(seed 1, frequency 4); 1920x1080 pinhole primaries from
//                 (0,0,-3) plus 4 cosine-weighted bounce rays per primary hit.
//   2  room:      262,144 triangles of axis-aligned boxes, 32-segment cylinders and thin slabs in
//                 a 30x12x18 room, triangle areas log-normal(mu=-6, sigma=2); 8,388,608 rays with
//                 origins uniform in the room and directions uniform on the sphere.
//   3  batch:     64 soup scenes, triangle counts log-uniform in [1e3, 2e5], ray counts
//                 log-uniform in [1e4, 2e6], uniform-sphere rays from interior points.
//   4  learned:   the room scene; training rays (stream 1) and fresh inference rays (stream 2).
//   5  bunny:     the reference validation camera (examples/validation/main.cpp:167-196),
//                 1024x1024 rays from (0,0,-0.5); the mesh is supplied by the caller.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

// SplitMix64 (Steele, Lea, Flood 2014).
struct Rng {
	uint64_t s;
	explicit Rng(uint64_t seed) : s(seed) {}
	uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
	// uniform in [0,1) with 53 random bits
	double u() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
	double range(double a, double b) { return a + (b - a) * u(); }
	double normal() {
		// Box-Muller; 1-u keeps the log argument in (0,1]
		const double u1 = 1.0 - u(), u2 = u();
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
	}
};

// Independent stream for element `index` of stream `stream` under `seed`.
inline Rng keyed(uint64_t seed, uint64_t stream, uint64_t index) {
	return Rng(mix64(mix64(seed * 0xD1B54A32D192ED03ull + stream) ^ (index * 0x9E3779B97F4A7C15ull)));
}

inline void put_tri(float *out, const double a[3], const double b[3], const double c[3]) {
	for (int k = 0; k < 3; ++k) {
		out[k] = static_cast<float>(a[k]);
		out[3 + k] = static_cast<float>(b[k]);
		out[6 + k] = static_cast<float>(c[k]);
	}
}

inline void put_ray(float *out, const double o[3], const double d[3]) {
	for (int k = 0; k < 3; ++k) {
		out[k] = static_cast<float>(o[k]);
		out[3 + k] = static_cast<float>(d[k]);
	}
}

inline void unit_sphere_dir(Rng &r, double d[3]) {
	const double z = 2.0 * r.u() - 1.0;
	const double phi = 2.0 * kPi * r.u();
	const double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
	d[0] = s * std::cos(phi);
	d[1] = s * std::sin(phi);
	d[2] = z;
}

// ---------------------------------------------------------------- soup (workloads 0 and 3)
void gen_soup(uint64_t seed, uint64_t n, float *out) {
	Rng r(mix64(seed ^ 0x50u));
	for (uint64_t i = 0; i < n; ++i) {
		double c[3], v[3][3];
		for (int k = 0; k < 3; ++k) c[k] = r.range(-1.0, 1.0);
		for (int j = 0; j < 3; ++j)
			for (int k = 0; k < 3; ++k) v[j][k] = c[k] + r.range(-0.05, 0.05);
		put_tri(out + 9 * i, v[0], v[1], v[2]);
	}
}

// Pinhole camera at `eye` looking along +z (towards the origin), vertical FOV 60 degrees,
// pixel (x, y) with y = 0 at the top row, ray index y * width + x.
void pinhole_ray(uint64_t idx, uint64_t width, uint64_t height, const double eye[3], float *out) {
	const uint64_t x = idx % width, y = idx / width;
	const double tan_half = std::tan(kPi / 6.0);
	const double aspect = static_cast<double>(width) / static_cast<double>(height);
	double d[3];
	d[0] = (2.0 * (static_cast<double>(x) + 0.5) / static_cast<double>(width) - 1.0) * tan_half * aspect;
	d[1] = (1.0 - 2.0 * (static_cast<double>(y) + 0.5) / static_cast<double>(height)) * tan_half;
	d[2] = 1.0;
	const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
	for (double &c : d) c /= len;
	put_ray(out, eye, d);
}

// ---------------------------------------------------------------- icosphere (workload 1)
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
	uint64_t h = mix64(seed * 0x632BE59BD9B4E019ull ^ static_cast<uint64_t>(x) * 0x9E3779B97F4A7C15ull);
	h = mix64(h ^ static_cast<uint64_t>(y) * 0xC2B2AE3D27D4EB4Full);
	h = mix64(h ^ static_cast<uint64_t>(z) * 0x165667B19E3779F9ull);
	return static_cast<double>(h >> 11) * 0x1.0p-52 - 1.0;  // [-1, 1)
}

double value_noise(const double p[3], uint64_t seed) {
	int64_t i[3];
	double f[3];
	for (int k = 0; k < 3; ++k) {
		const double fl = std::floor(p[k]);
		i[k] = static_cast<int64_t>(fl);
		const double t = p[k] - fl;
		f[k] = t * t * (3.0 - 2.0 * t);
	}
	double acc = 0.0;
	for (int c = 0; c < 8; ++c) {
		const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
		const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
		acc += w * lattice(i[0] + dx, i[1] + dy, i[2] + dz, seed);
	}
	return acc;
}

// Icosphere subdivided `levels` times (20 * 4^levels triangles), vertices at
// centre + radius * (1 + amp * value_noise(freq * unit position)) along the unit direction.
void gen_icosphere(uint64_t seed, int levels, const double centre[3], double radius, double amp, double freq,
                   float *out) {
	const double t = (1.0 + std::sqrt(5.0)) / 2.0;
	std::vector<std::array<double, 3>> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
	                                        {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
	                                        {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
	for (auto &p : v) {
		const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
		for (double &c : p) c /= l;
	}
	std::vector<std::array<uint32_t, 3>> f = {
	    {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
	    {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
	    {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
	for (int level = 0; level < levels; ++level) {
		std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
		auto midpoint = [&](uint32_t a, uint32_t b) {
			const auto key = std::make_pair(a < b ? a : b, a < b ? b : a);
			auto it = mid.find(key);
			if (it != mid.end()) return it->second;
			std::array<double, 3> m;
			for (int k = 0; k < 3; ++k) m[k] = 0.5 * (v[a][k] + v[b][k]);
			const double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
			for (double &c : m) c /= l;
			v.push_back(m);
			const uint32_t id = static_cast<uint32_t>(v.size() - 1);
			mid.emplace(key, id);
			return id;
		};
		std::vector<std::array<uint32_t, 3>> nf;
		nf.reserve(f.size() * 4);
		for (const auto &tri : f) {
			const uint32_t ab = midpoint(tri[0], tri[1]), bc = midpoint(tri[1], tri[2]),
			               ca = midpoint(tri[2], tri[0]);
			nf.push_back({tri[0], ab, ca});
			nf.push_back({tri[1], bc, ab});
			nf.push_back({tri[2], ca, bc});
			nf.push_back({ab, bc, ca});
		}
		f.swap(nf);
	}
	std::vector<std::array<double, 3>> disp(v.size());
	for (size_t i = 0; i < v.size(); ++i) {
		const double q[3] = {freq * v[i][0], freq * v[i][1], freq * v[i][2]};
		const double r = radius * (1.0 + amp * value_noise(q, seed));
		for (int k = 0; k < 3; ++k) disp[i][k] = centre[k] + r * v[i][k];
	}
	for (size_t i = 0; i < f.size(); ++i)
		put_tri(out + 9 * i, disp[f[i][0]].data(), disp[f[i][1]].data(), disp[f[i][2]].data());
}

// The displaced unit icosphere of workload 1: 6 subdivisions, 81,920 triangles.
void gen_sphere_w1(float *out) {
	const double c[3] = {0.0, 0.0, 0.0};
	gen_icosphere(1, 6, c, 1.0, 0.08, 4.0, out);
}

// A closed stand-in for the reference's validation mesh (examples/common/bunny.obj, which has no
// licence file and stays out of this repository), used with its cameras (workloads 5 and 6):
// an icosphere subdivided 4 times (5,120 triangles; the bunny has 4,968) centred in the bunny's
// bounding box ([-0.094, 0.061] x [0.033, 0.187] x [-0.062, 0.059]), radius 0.07, displaced by
// 0.25 * value noise (seed 5, frequency 3).
constexpr uint64_t kBlobTris = 5120;
void gen_blob(float *out) {
	const double c[3] = {-0.017, 0.11, -0.0015};
	gen_icosphere(5, 4, c, 0.07, 0.25, 3.0, out);
}

// ---------------------------------------------------------------- room (workloads 2 and 4)
constexpr double kRoom[3] = {30.0, 12.0, 18.0};

struct TriSink {
	float *out;
	uint64_t n, cap;
	bool full() const { return n >= cap; }
	void add(const double a[3], const double b[3], const double c[3]) {
		if (n < cap) put_tri(out + 9 * n, a, b, c);
		++n;
	}
	void quad(const double a[3], const double b[3], const double c[3], const double d[3]) {
		add(a, b, c);
		add(a, c, d);
	}
};

void emit_box(TriSink &s, const double lo[3], const double hi[3]) {
	double p[8][3];
	for (int c = 0; c < 8; ++c)
		for (int k = 0; k < 3; ++k) p[c][k] = ((c >> k) & 1) ? hi[k] : lo[k];
	// faces -x +x -y +y -z +z (corner index bit k = axis k at hi)
	s.quad(p[0], p[4], p[6], p[2]);
	s.quad(p[1], p[3], p[7], p[5]);
	s.quad(p[0], p[1], p[5], p[4]);
	s.quad(p[2], p[6], p[7], p[3]);
	s.quad(p[0], p[2], p[3], p[1]);
	s.quad(p[4], p[5], p[7], p[6]);
}

void emit_cylinder(TriSink &s, const double base[3], double radius, double height) {
	constexpr int kSeg = 32;
	double ring[kSeg][2];
	for (int i = 0; i < kSeg; ++i) {
		const double a = 2.0 * kPi * static_cast<double>(i) / kSeg;
		ring[i][0] = std::cos(a);
		ring[i][1] = std::sin(a);
	}
	const double cb[3] = {base[0], base[1], base[2]};
	const double ct[3] = {base[0], base[1] + height, base[2]};
	for (int i = 0; i < kSeg; ++i) {
		const int j = (i + 1) % kSeg;
		const double b0[3] = {base[0] + radius * ring[i][0], base[1], base[2] + radius * ring[i][1]};
		const double b1[3] = {base[0] + radius * ring[j][0], base[1], base[2] + radius * ring[j][1]};
		const double t0[3] = {b0[0], base[1] + height, b0[2]};
		const double t1[3] = {b1[0], base[1] + height, b1[2]};
		s.quad(b0, b1, t1, t0);
		s.add(cb, b1, b0);
		s.add(ct, t0, t1);
	}
}

void gen_room(uint64_t seed, uint64_t n, float *out) {
	TriSink s{out, 0, n};
	const double lo[3] = {0, 0, 0};
	// six walls
	{
		double p[8][3];
		for (int c = 0; c < 8; ++c)
			for (int k = 0; k < 3; ++k) p[c][k] = ((c >> k) & 1) ? kRoom[k] : lo[k];
		s.quad(p[0], p[2], p[6], p[4]);
		s.quad(p[1], p[5], p[7], p[3]);
		s.quad(p[0], p[4], p[5], p[1]);
		s.quad(p[2], p[3], p[7], p[6]);
		s.quad(p[0], p[1], p[3], p[2]);
		s.quad(p[4], p[6], p[7], p[5]);
	}
	Rng r(mix64(seed ^ 0x2000u));
	const double sin_half_seg = std::sin(kPi / 32.0);
	while (!s.full()) {
		const double kind = r.u();
		// target triangle area, log-normal(mu=-6, sigma=2), clamped so objects fit the room
		double area = std::exp(-6.0 + 2.0 * r.normal());
		area = std::fmin(std::fmax(area, 1e-7), 8.0);
		if (kind < 0.45) {  // axis-aligned box: face triangles have area ~ side^2 / 2
			const double side = std::sqrt(2.0 * area);
			double dim[3], mn[3], mx[3];
			for (int k = 0; k < 3; ++k) dim[k] = std::fmin(side * r.range(0.5, 2.0), 0.9 * kRoom[k]);
			for (int k = 0; k < 3; ++k) {
				mn[k] = r.range(0.0, kRoom[k] - dim[k]);
				mx[k] = mn[k] + dim[k];
			}
			emit_box(s, mn, mx);
		} else if (kind < 0.75) {  // thin slab: two large faces carry the area
			const int thin = static_cast<int>(r.next() % 3);
			const double side = std::sqrt(2.0 * area);
			double dim[3], mn[3], mx[3];
			for (int k = 0; k < 3; ++k) dim[k] = side * r.range(0.5, 2.0);
			dim[thin] = 0.02 * std::fmin(dim[(thin + 1) % 3], dim[(thin + 2) % 3]);
			for (int k = 0; k < 3; ++k) dim[k] = std::fmin(dim[k], 0.9 * kRoom[k]);
			for (int k = 0; k < 3; ++k) {
				mn[k] = r.range(0.0, kRoom[k] - dim[k]);
				mx[k] = mn[k] + dim[k];
			}
			emit_box(s, mn, mx);
		} else {  // vertical cylinder: side triangles have area chord * height / 2
			const double aspect = r.range(0.5, 4.0);  // height / radius
			double radius = std::sqrt(area / (sin_half_seg * aspect));
			radius = std::fmin(radius, 0.45 * std::fmin(kRoom[0], kRoom[2]));
			const double height = std::fmin(aspect * radius, 0.9 * kRoom[1]);
			const double base[3] = {r.range(radius, kRoom[0] - radius), r.range(0.0, kRoom[1] - height),
			                        r.range(radius, kRoom[2] - radius)};
			emit_cylinder(s, base, radius, height);
		}
	}
}

void room_ray(uint64_t seed, uint64_t stream, uint64_t idx, float *out) {
	Rng r = keyed(seed, stream, idx);
	double o[3], d[3];
	for (int k = 0; k < 3; ++k) o[k] = kRoom[k] * r.u();
	unit_sphere_dir(r, d);
	put_ray(out, o, d);
}

// ---------------------------------------------------------------- batch (workload 3)
constexpr int kBatchScenes = 64;

uint64_t log_uniform_count(uint64_t seed, uint64_t stream, uint64_t idx, double lo, double hi) {
	Rng r = keyed(seed, stream, idx);
	return static_cast<uint64_t>(std::llround(std::exp(r.range(std::log(lo), std::log(hi)))));
}

uint64_t batch_tris(uint64_t i) { return log_uniform_count(3, 10, i, 1e3, 2e5); }
uint64_t batch_rays(uint64_t i) { return log_uniform_count(3, 11, i, 1e4, 2e6); }

void soup_interior_ray(uint64_t seed, uint64_t stream, uint64_t idx, float *out) {
	Rng r = keyed(seed, stream, idx);
	double o[3], d[3];
	for (double &c : o) c = r.range(-1.0, 1.0);
	unit_sphere_dir(r, d);
	put_ray(out, o, d);
}

// ---------------------------------------------------------------- bunny camera (workload 5)
// Restates the camera of examples/validation/main.cpp:167-196: float arithmetic as written
// there, directions divided by the length of the sensor position (not unit length).
void bunny_ray(uint64_t idx, float *out) {
	const int width = 1024, height = 1024;
	const int row = static_cast<int>(idx / width);
	const int x = static_cast<int>(idx % width);
	const int y = height - 1 - row;
	const float camera_dist = 0.5f, sensor_size = 0.05f, sensor_dist = 0.05f;
	const float sx = static_cast<float>(sensor_size * (static_cast<float>(x) / width - 0.5));
	const float sy = static_cast<float>(sensor_size * (static_cast<float>(y) / height - 0.5));
	const float cam[3] = {0.0f, 0.0f, -camera_dist};
	const float sensor[3] = {sx, sy, -camera_dist + sensor_dist};
	float d[3] = {sensor[0] - cam[0], sensor[1] - cam[1], sensor[2] - cam[2]};
	const float len = static_cast<float>(
	    std::sqrt(sensor[0] * sensor[0] + sensor[1] * sensor[1] + sensor[2] * sensor[2]));
	for (float &c : d) c /= len;
	for (int k = 0; k < 3; ++k) {
		out[k] = cam[k];
		out[3 + k] = d[k];
	}
}

// The camera of the bunny benchmark (examples/bunny/main.cpp:84-122): 4096 x 4096 pixels, eye
// at (0, -0.02, -0.3), sensor 0.05 wide at distance 0.05, directions of unit length (divided by
// their own length, in float as written there), rows from the top.
void bunny_bench_ray(uint64_t idx, float *out) {
	const int width = 4096, height = 4096;
	const int row = static_cast<int>(idx / width);
	const int x = static_cast<int>(idx % width);
	const int y = height - 1 - row;
	const float camera_dist = 0.3f, sensor_size = 0.05f, sensor_dist = 0.05f;
	const float sx = static_cast<float>(sensor_size * (static_cast<float>(x) / width - 0.5));
	const float sy = static_cast<float>(sensor_size * (static_cast<float>(y) / height - 0.5));
	const float cam[3] = {0.0f, -0.02f, -camera_dist};
	const float sensor[3] = {sx, sy, -camera_dist + sensor_dist};
	float d[3] = {sensor[0] - cam[0], sensor[1] - cam[1], sensor[2] - cam[2]};
	const float len = static_cast<float>(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
	for (float &c : d) c /= len;
	for (int k = 0; k < 3; ++k) {
		out[k] = cam[k];
		out[3 + k] = d[k];
	}
}

// Cosine-weighted hemisphere direction about unit normal n (branchless orthonormal basis of
// Duff et al. 2017).
void cosine_dir(Rng &r, const double n[3], double d[3]) {
	const double u1 = r.u(), u2 = r.u();
	const double rad = std::sqrt(u1), phi = 2.0 * kPi * u2;
	const double lx = rad * std::cos(phi), ly = rad * std::sin(phi), lz = std::sqrt(std::fmax(0.0, 1.0 - u1));
	const double sign = std::copysign(1.0, n[2]);
	const double a = -1.0 / (sign + n[2]);
	const double b = n[0] * n[1] * a;
	const double t[3] = {1.0 + sign * n[0] * n[0] * a, sign * b, -sign * n[0]};
	const double bt[3] = {b, sign + n[1] * n[1] * a, -n[1]};
	for (int k = 0; k < 3; ++k) d[k] = lx * t[k] + ly * bt[k] + lz * n[k];
	const double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
	for (int k = 0; k < 3; ++k) d[k] /= l;
}

}  // namespace

extern "C" {

// Number of scenes in a workload (64 for the batch, otherwise 1).
int prt_workload_scenes(int workload) { return workload == 3 ? kBatchScenes : 1; }

// Triangle count of scene `scene` of `workload`; 0 for an unknown workload.
uint64_t prt_scene_tris(int workload, int scene) {
	switch (workload) {
	case 0: return 16384;
	case 1: return 81920;
	case 2:
	case 4: return 262144;
	case 3: return (scene >= 0 && scene < kBatchScenes) ? batch_tris(static_cast<uint64_t>(scene)) : 0;
	case 5:
	case 6: return kBlobTris;
	default: return 0;
	}
}

// Writes prt_scene_tris(workload, scene) triangles (9 floats each). Returns 0 on success.
int prt_scene_generate(int workload, int scene, float *tris) {
	switch (workload) {
	case 0: gen_soup(0, 16384, tris); return 0;
	case 1: gen_sphere_w1(tris); return 0;
	case 5:
	case 6: gen_blob(tris); return 0;
	case 2:
	case 4: gen_room(2, 262144, tris); return 0;
	case 3:
		if (scene < 0 || scene >= kBatchScenes) return 1;
		gen_soup(mix64(3 * 1000003ull + static_cast<uint64_t>(scene)), batch_tris(scene), tris);
		return 0;
	default: return 1;
	}
}

// Ray count of the generated (not traced) ray set: primaries for workload 1, stream 1
// (training) or 2 (inference) for workload 4.
uint64_t prt_rays_count(int workload, int scene, int stream) {
	switch (workload) {
	case 0: return 1024ull * 1024ull;
	case 1: return 1920ull * 1080ull;
	case 2: return 8388608ull;
	case 3: return (scene >= 0 && scene < kBatchScenes) ? batch_rays(static_cast<uint64_t>(scene)) : 0;
	case 4: return stream == 1 ? 16777216ull : 8388608ull;
	case 5: return 1024ull * 1024ull;
	case 6: return 4096ull * 4096ull;
	default: return 0;
	}
}

// Writes rays [first, first + count) of the ray set (6 floats each). Returns 0 on success.
int prt_rays_generate(int workload, int scene, int stream, uint64_t first, uint64_t count, float *rays) {
	const uint64_t total = prt_rays_count(workload, scene, stream);
	if (first + count > total || first + count < first) return 1;
	const double eye[3] = {0.0, 0.0, -3.0};
	for (uint64_t i = 0; i < count; ++i) {
		const uint64_t idx = first + i;
		float *o = rays + 6 * i;
		switch (workload) {
		case 0: pinhole_ray(idx, 1024, 1024, eye, o); break;
		case 1: pinhole_ray(idx, 1920, 1080, eye, o); break;
		case 2: room_ray(2, 1, idx, o); break;
		case 3: soup_interior_ray(mix64(3 * 7919ull + static_cast<uint64_t>(scene)), 1, idx, o); break;
		case 4: room_ray(4, static_cast<uint64_t>(stream), idx, o); break;
		case 5: bunny_ray(idx, o); break;
		case 6: bunny_bench_ray(idx, o); break;
		default: return 1;
		}
	}
	return 0;
}

// Bounce rays for workload 1: for primary ray i with a finite hit (t, primitive id, hit point p
// as returned by nearest_hits), 4 cosine-weighted rays about the geometric normal facing the
// incoming ray, starting at p + 1e-4 * normal. Rays of misses are skipped. `out` holds room for
// 4 * n rays; the number written is returned in *n_out. Stream keyed by (seed 1, i, k).
int prt_bounce_rays(const float *tris, uint64_t n_tris, const float *prim_rays, const float *t,
                    const uint32_t *prim_id, const float *p, uint64_t n, float *out, uint64_t *n_out) {
	uint64_t w = 0;
	for (uint64_t i = 0; i < n; ++i) {
		if (!(t[i] < INFINITY) || prim_id[i] >= n_tris) continue;
		const float *v = tris + 9ull * prim_id[i];
		double e1[3], e2[3], nrm[3];
		for (int k = 0; k < 3; ++k) {
			e1[k] = static_cast<double>(v[3 + k]) - v[k];
			e2[k] = static_cast<double>(v[6 + k]) - v[k];
		}
		nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
		nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
		nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
		const double l = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
		if (!(l > 0.0)) continue;
		for (double &c : nrm) c /= l;
		const float *d = prim_rays + 6 * i + 3;
		if (nrm[0] * d[0] + nrm[1] * d[1] + nrm[2] * d[2] > 0.0)
			for (double &c : nrm) c = -c;
		const double org[3] = {p[3 * i] + 1e-4 * nrm[0], p[3 * i + 1] + 1e-4 * nrm[1], p[3 * i + 2] + 1e-4 * nrm[2]};
		for (uint64_t k = 0; k < 4; ++k) {
			Rng r = keyed(1, 20, i * 4 + k);
			double dir[3];
			cosine_dir(r, nrm, dir);
			put_ray(out + 6 * w, org, dir);
			++w;
		}
	}
	*n_out = w;
	return 0;
}

}  // extern "C"
