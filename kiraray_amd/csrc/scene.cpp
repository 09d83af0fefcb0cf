// scene.cpp - procedural scenes behind krr_scene_create (host only).
//
// Stands in for SceneImporter::import (src/scene/krrscene.cpp:253-305): a
// JSON document lists "model" entries, "materials", "media", "camera" and
// "environment".  Models are generated procedurally instead of loaded from
// OBJ/glTF files:
//   {"type": "cornell_box"}
//   {"type": "icosphere_field", "count", "level", "seed", "radius": [lo,hi],
//    "fractions": [diffuse, conductor], "ground": bool, "lights": n}
//   {"type": "quad", "corners": [[x,y,z] x4], "material": i, "emission": [r,g,b]}
//   {"type": "mesh", "positions": [...], "indices": [...], "normals": [...],
//    "material": i, "emission": [r,g,b]}
// Media: {"type": "value_noise_grid", "res", "octaves", "seed",
//         "sigma_t_max", "albedo", "g", "threshold", "bounds": [[lo],[hi]]}
//         (density = max(0, noise - threshold) scaled to max 1; threshold 0)
//        {"type": "homogeneous", "sigma_t", "albedo", "g", "Le"}
//        {"type": "dense_grid", "res", "density", "bounds", "sigma_t",
//         "albedo", "g", "scale"}
// Camera: {"eye", "target", "up", "focalLength", "focalDistance",
//          "lensRadius", "filmSize", "medium"} with the aspect ratio taken
// from "resolution" as the renderer does (src/main/renderer.cpp:28).
// Views: {"count", "seed", "radius", "center", "sizes": [...]} creates many
// cameras on a sphere looking at "center", each with a square resolution
// drawn from "sizes".
//
// KiRaRay's own scene JSON (SceneImporter::import, krrscene.cpp:253-350) is
// read as well.  "model" is a node tree (arrays, file paths, objects with
// "name", "translate", "rotate" [w,x,y,z], "scale", "params" and a nested
// "model"); transforms compose down the tree.  Wavefront OBJ + MTL files
// are loaded (paths relative to "base_dir", default the working directory);
// "type": "medium" nodes add homogeneous media with the reference's keys,
// a "bound" box and mesh bindings by name; "type": "light" nodes add the
// constant infinite light.  "materials" entries {"name", "bsdf", "params"}
// override the file's materials of that name (krrscene.cpp:160-238), and
// "camera": {"mData": {...}} with "cameraController": {"mData": {"target",
// "radius", "pitch", "yaw"}} places the camera as the orbit controller does
// (camera.cpp:48-61).  OpenVDB float grids (.vdb model files and
// "heterogeneous" media with a "file") become dense-grid media inside a
// null-material box (OpenVDBImporter::import, openvdb.cpp:8-67).  What this
// build cannot render is an error: other model formats, temperature and
// albedo grids, point / spot / directional lights.
#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <memory>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "errors.h"
#include "json.h"
#include "krr_wavefront.h"

struct krr_scene {
	std::vector<float> positions, normals, texcoords, tangents;
	std::vector<int32_t> indices;
	std::vector<krr_mesh> meshes;
	std::vector<std::string> meshNames; // one per mesh ("" when unnamed)
	std::vector<krr_material> materials;
	std::vector<krr_medium> media;
	std::vector<std::unique_ptr<std::vector<float>>> densities;
	std::vector<krr_infinite_light> infinite;
	std::vector<std::unique_ptr<std::vector<float>>> images; // RGBA, held for infinite lights and textures
	std::vector<krr_texture> textures;
	std::map<std::string, int32_t> textureByPath; // 1 + index into textures
	std::vector<krr_camera> cameras;
	std::vector<std::pair<int32_t, int32_t>> viewSizes;
	bool anyNormals = false, anyTexcoords = false, anyTangents = false;
};

namespace krr {

namespace {

// Small PCG32 for scene randomness (seeded per purpose).
struct Rng {
	uint64_t state, inc;
	Rng(uint64_t seed, uint64_t stream) {
		state = 0;
		inc	  = (stream << 1u) | 1u;
		next();
		state += seed;
		next();
	}
	uint32_t next() {
		uint64_t old = state;
		state		 = old * 0x5851f42d4c957f2dULL + inc;
		uint32_t xs	 = (uint32_t) (((old >> 18u) ^ old) >> 27u);
		uint32_t rot = (uint32_t) (old >> 59u);
		return (xs >> rot) | (xs << ((~rot + 1u) & 31));
	}
	float uniform() { return (float) (next() >> 8) * (1.0f / 16777216.0f); }
	float range(float lo, float hi) { return lo + (hi - lo) * uniform(); }
};

struct V { float x, y, z; };
inline V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V add(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V mul(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V crs(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dt(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V nrmz(V a) {
	float l = std::sqrt(dt(a, a));
	return l > 0 ? mul(a, 1.f / l) : a;
}

krr_material makeMaterial(int type, V diffuse, float roughness, float ior, float specTrans) {
	krr_material m;
	std::memset(&m, 0, sizeof(m));
	m.type		  = type;
	m.diffuse[0]  = diffuse.x;
	m.diffuse[1]  = diffuse.y;
	m.diffuse[2]  = diffuse.z;
	m.specular[3] = 1.f - roughness; // SpecularGlossiness: a = 1 - roughness (krrscene.cpp:207)
	m.ior		  = ior;
	m.specular_transmission = specTrans;
	return m;
}

struct Builder {
	krr_scene &s;
	int addMaterial(const krr_material &m) {
		s.materials.push_back(m);
		return (int) s.materials.size() - 1;
	}
	// Appends a mesh; normals, texture coordinates and tangents may be empty.
	int addMesh(const std::vector<V> &verts, const std::vector<V> &nrms, const std::vector<int> &tris, int material,
				V emission = {0, 0, 0}, int mIn = -1, int mOut = -1, const std::string &name = "",
				const std::vector<float> &uvs = {}, const std::vector<V> &tans = {}) {
		krr_mesh m;
		std::memset(&m, 0, sizeof(m));
		int32_t base	 = (int32_t) (s.positions.size() / 3);
		m.first_triangle = (int32_t) (s.indices.size() / 3);
		m.n_triangles	 = (int32_t) (tris.size() / 3);
		m.material		 = material;
		m.medium_inside	 = mIn;
		m.medium_outside = mOut;
		m.has_normals	 = nrms.empty() ? 0 : 1;
		m.has_texcoords	 = uvs.empty() ? 0 : 1;
		m.has_tangents	 = tans.empty() ? 0 : 1;
		m.emission[0]	 = emission.x;
		m.emission[1]	 = emission.y;
		m.emission[2]	 = emission.z;
		for (size_t i = 0; i < verts.size(); i++) {
			s.positions.push_back(verts[i].x);
			s.positions.push_back(verts[i].y);
			s.positions.push_back(verts[i].z);
			V n = nrms.empty() ? V{0, 0, 0} : nrms[i];
			s.normals.push_back(n.x);
			s.normals.push_back(n.y);
			s.normals.push_back(n.z);
			s.texcoords.push_back(uvs.empty() ? 0.f : uvs[2 * i]);
			s.texcoords.push_back(uvs.empty() ? 0.f : uvs[2 * i + 1]);
			V t = tans.empty() ? V{0, 0, 0} : tans[i];
			s.tangents.push_back(t.x);
			s.tangents.push_back(t.y);
			s.tangents.push_back(t.z);
		}
		if (!nrms.empty()) s.anyNormals = true;
		if (!uvs.empty()) s.anyTexcoords = true;
		if (!tans.empty()) s.anyTangents = true;
		for (int t : tris) s.indices.push_back(base + t);
		s.meshes.push_back(m);
		s.meshNames.push_back(name);
		return (int) s.meshes.size() - 1;
	}
	// Quad p0 p1 p2 p3 (counter-clockwise seen from the side its normal faces).
	int addQuad(V p0, V p1, V p2, V p3, int material, V emission = {0, 0, 0}, bool withNormals = false) {
		std::vector<V> v = {p0, p1, p2, p3};
		std::vector<V> n;
		if (withNormals) {
			V fn = nrmz(crs(sub(p1, p0), sub(p2, p0)));
			n	 = {fn, fn, fn, fn};
		}
		return addMesh(v, n, {0, 1, 2, 0, 2, 3}, material, emission);
	}
	// Axis-aligned box [lo,hi] rotated about the vertical axis through its centre.
	int addBox(V lo, V hi, float angleDeg, int material) {
		V c		= mul(add(lo, hi), 0.5f);
		float a = angleDeg * 3.14159265358979f / 180.f;
		float ca = std::cos(a), sa = std::sin(a);
		auto rot = [&](V p) {
			V q = sub(p, c);
			return V{c.x + ca * q.x + sa * q.z, p.y, c.z - sa * q.x + ca * q.z};
		};
		std::vector<V> v;
		for (int i = 0; i < 8; i++) v.push_back(rot({(i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z}));
		// outward counter-clockwise faces
		std::vector<int> t = {0, 4, 6, 0, 6, 2,  1, 3, 7, 1, 7, 5,  0, 1, 5, 0, 5, 4,
							  2, 6, 7, 2, 7, 3,  0, 2, 3, 0, 3, 1,  4, 5, 7, 4, 7, 6};
		return addMesh(v, {}, t, material);
	}
};

// Cornell-style box: floor, ceiling, back, left and right walls (10 tris),
// two rotated boxes (24 tris) and a ceiling light quad (2 tris) -> 36.
void cornellBox(Builder &b, const Json &) {
	int white = b.addMaterial(makeMaterial(KRR_MAT_DIFFUSE, {0.73f, 0.73f, 0.73f}, 1.f, 1.5f, 0.f));
	int red	  = b.addMaterial(makeMaterial(KRR_MAT_DIFFUSE, {0.63f, 0.065f, 0.05f}, 1.f, 1.5f, 0.f));
	int green = b.addMaterial(makeMaterial(KRR_MAT_DIFFUSE, {0.14f, 0.45f, 0.09f}, 1.f, 1.5f, 0.f));
	int lightMat = b.addMaterial(makeMaterial(KRR_MAT_DIFFUSE, {0.78f, 0.78f, 0.78f}, 1.f, 1.5f, 0.f));
	const float L = -1, R = 1, B = 0, T = 2, Bk = -1, F = 1;
	b.addQuad({L, B, F}, {R, B, F}, {R, B, Bk}, {L, B, Bk}, white);		 // floor (+y)
	b.addQuad({L, T, Bk}, {R, T, Bk}, {R, T, F}, {L, T, F}, white);		 // ceiling (-y)
	b.addQuad({L, B, Bk}, {R, B, Bk}, {R, T, Bk}, {L, T, Bk}, white);	 // back (+z)
	b.addQuad({L, B, F}, {L, B, Bk}, {L, T, Bk}, {L, T, F}, red);		 // left (+x)
	b.addQuad({R, B, Bk}, {R, B, F}, {R, T, F}, {R, T, Bk}, green);		 // right (-x)
	b.addBox({0.05f, 0.f, -0.05f}, {0.65f, 0.6f, 0.55f}, -18.f, white);	 // short box
	b.addBox({-0.7f, 0.f, -0.75f}, {-0.1f, 1.2f, -0.15f}, 15.f, white);	 // tall box
	const float h = 1.98f, e = 0.25f;
	b.addQuad({-e, h, -e}, {e, h, -e}, {e, h, e}, {-e, h, e}, lightMat, {15.f, 15.f, 15.f}, true); // faces -y
}

// Icosphere with `level` subdivisions (20 * 4^level triangles).
void icosphere(int level, std::vector<V> &verts, std::vector<int> &tris) {
	const float t = (1.f + std::sqrt(5.f)) / 2.f;
	verts		  = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
					 {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
	for (auto &v : verts) v = nrmz(v);
	tris = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
			3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
	for (int l = 0; l < level; l++) {
		std::map<uint64_t, int> mid;
		auto midpoint = [&](int a, int c) {
			uint64_t key = a < c ? ((uint64_t) a << 32 | (uint32_t) c) : ((uint64_t) c << 32 | (uint32_t) a);
			auto it		 = mid.find(key);
			if (it != mid.end()) return it->second;
			V m = nrmz(mul(add(verts[a], verts[c]), 0.5f));
			verts.push_back(m);
			int idx	 = (int) verts.size() - 1;
			mid[key] = idx;
			return idx;
		};
		std::vector<int> nt;
		nt.reserve(tris.size() * 4);
		for (size_t i = 0; i < tris.size(); i += 3) {
			int a = tris[i], bb = tris[i + 1], c = tris[i + 2];
			int ab = midpoint(a, bb), bc = midpoint(bb, c), ca = midpoint(c, a);
			int add4[12] = {a, ab, ca, bb, bc, ab, c, ca, bc, ab, bc, ca};
			nt.insert(nt.end(), add4, add4 + 12);
		}
		tris.swap(nt);
	}
}

// 1024 icospheres in the unit cube over a ground quad with 16 quad lights.
void icosphereField(Builder &b, const Json &j) {
	int count		= (int) j.value("count", 1024.0);
	int level		= (int) j.value("level", 3.0);
	uint64_t seed	= (uint64_t) j.value("seed", 7.0);
	float rad[2]	= {0.01f, 0.05f};
	j.vec<2>("radius", rad);
	float frac[2] = {0.5f, 0.3f};
	j.vec<2>("fractions", frac);
	bool ground = j.value("ground", true);
	int nLights = (int) j.value("lights", 16.0);
	float Le	= (float) j.value("light_radiance", 12.0);
	if (count < 0 || level < 0 || level > 6) throw std::runtime_error("icosphere_field: bad count/level");
	std::vector<V> unitV;
	std::vector<int> unitT;
	icosphere(level, unitV, unitT);
	Rng rng(seed, 1);
	for (int s = 0; s < count; s++) {
		V c		= {rng.uniform(), rng.uniform(), rng.uniform()};
		float r = rng.range(rad[0], rad[1]);
		float u = rng.uniform();
		krr_material m;
		if (u < frac[0]) m = makeMaterial(KRR_MAT_DIFFUSE, {rng.range(0.1f, 0.9f), rng.range(0.1f, 0.9f), rng.range(0.1f, 0.9f)}, 1.f, 1.5f, 0.f);
		else if (u < frac[0] + frac[1]) {
			V refl	= {rng.range(0.5f, 0.95f), rng.range(0.5f, 0.95f), rng.range(0.5f, 0.95f)};
			m		= makeMaterial(KRR_MAT_CONDUCTOR, refl, rng.range(0.05f, 0.5f), 1.5f, 0.f);
		} else m = makeMaterial(KRR_MAT_DIELECTRIC, {1.f, 1.f, 1.f}, 0.f, 1.5f, 1.f);
		int mat = b.addMaterial(m);
		std::vector<V> verts(unitV.size()), nrms(unitV.size());
		for (size_t i = 0; i < unitV.size(); i++) {
			verts[i] = add(c, mul(unitV[i], r));
			nrms[i]	 = unitV[i];
		}
		b.addMesh(verts, nrms, unitT, mat);
	}
	if (ground) {
		int g = b.addMaterial(makeMaterial(KRR_MAT_DIFFUSE, {0.5f, 0.5f, 0.5f}, 1.f, 1.5f, 0.f));
		const float lo = -8.f, hi = 9.f;
		b.addQuad({lo, 0, hi}, {hi, 0, hi}, {hi, 0, lo}, {lo, 0, lo}, g);
	}
	if (nLights > 0) {
		int lm	= b.addMaterial(makeMaterial(KRR_MAT_DIFFUSE, {0.8f, 0.8f, 0.8f}, 1.f, 1.5f, 0.f));
		int n	= (int) std::ceil(std::sqrt((double) nLights));
		float h = 1.8f, e = 0.075f;
		for (int i = 0; i < nLights; i++) {
			float cx = (i % n + 0.5f) / n, cz = (i / n + 0.5f) / n;
			b.addQuad({cx - e, h, cz - e}, {cx + e, h, cz - e}, {cx + e, h, cz + e}, {cx - e, h, cz + e}, lm, {Le, Le, Le}, true);
		}
	}
}

void explicitMesh(Builder &b, const Json &j) {
	const auto &P = j.at("positions").items();
	const auto &I = j.at("indices").items();
	if (P.size() % 3 || I.size() % 3) throw std::runtime_error("mesh: positions/indices must be multiples of 3");
	std::vector<V> verts, nrms;
	for (size_t i = 0; i < P.size(); i += 3) verts.push_back({(float) P[i].asNumber(), (float) P[i + 1].asNumber(), (float) P[i + 2].asNumber()});
	if (j.contains("normals")) {
		const auto &N = j.at("normals").items();
		if (N.size() != P.size()) throw std::runtime_error("mesh: normals must match positions");
		for (size_t i = 0; i < N.size(); i += 3) nrms.push_back({(float) N[i].asNumber(), (float) N[i + 1].asNumber(), (float) N[i + 2].asNumber()});
	}
	std::vector<int> tris;
	for (auto &x : I) {
		int v = (int) x.asNumber();
		if (v < 0 || (size_t) v >= verts.size()) throw std::runtime_error("mesh: index out of range");
		tris.push_back(v);
	}
	std::vector<float> uvs;
	if (j.contains("texcoords")) {
		const auto &T = j.at("texcoords").items();
		if (T.size() != 2 * verts.size()) throw std::runtime_error("mesh: texcoords must hold 2 values per vertex");
		for (const Json &t : T) uvs.push_back((float) t.asNumber());
	}
	std::vector<V> tans;
	if (j.contains("tangents")) {
		const auto &T = j.at("tangents").items();
		if (T.size() != P.size()) throw std::runtime_error("mesh: tangents must match positions");
		for (size_t i = 0; i < T.size(); i += 3) tans.push_back({(float) T[i].asNumber(), (float) T[i + 1].asNumber(), (float) T[i + 2].asNumber()});
	}
	float em[3] = {0, 0, 0};
	j.vec<3>("emission", em);
	int mat	 = (int) j.value("material", 0.0);
	int mIn	 = (int) j.value("medium_inside", -1.0);
	int mOut = (int) j.value("medium_outside", -1.0);
	b.addMesh(verts, nrms, tris, mat, {em[0], em[1], em[2]}, mIn, mOut, j.value("name", ""), uvs, tans);
}

void quadModel(Builder &b, const Json &j) {
	const auto &C = j.at("corners").items();
	if (C.size() != 4) throw std::runtime_error("quad: corners must hold 4 points");
	V p[4];
	for (int i = 0; i < 4; i++) {
		const auto &c = C[i].items();
		if (c.size() != 3) throw std::runtime_error("quad: corner must be [x,y,z]");
		p[i] = {(float) c[0].asNumber(), (float) c[1].asNumber(), (float) c[2].asNumber()};
	}
	float em[3] = {0, 0, 0};
	j.vec<3>("emission", em);
	bool emissive = em[0] > 0 || em[1] > 0 || em[2] > 0;
	b.addQuad(p[0], p[1], p[2], p[3], (int) j.value("material", 0.0), {em[0], em[1], em[2]}, emissive || j.value("normals", false));
}

// Hash of an integer lattice point to [0,1).
inline float latticeValue(int x, int y, int z, uint32_t salt) {
	uint32_t h = (uint32_t) x * 0x8da6b343u ^ (uint32_t) y * 0xd8163841u ^ (uint32_t) z * 0xcb1ab31fu ^ salt * 0x9e3779b9u;
	h ^= h >> 16;
	h *= 0x7feb352du;
	h ^= h >> 15;
	h *= 0x846ca68bu;
	h ^= h >> 16;
	return (float) (h >> 8) * (1.0f / 16777216.0f);
}

inline float valueNoise(float x, float y, float z, uint32_t salt) {
	int ix = (int) std::floor(x), iy = (int) std::floor(y), iz = (int) std::floor(z);
	float fx = x - ix, fy = y - iy, fz = z - iz;
	auto sm = [](float t) { return t * t * (3.f - 2.f * t); };
	fx = sm(fx);
	fy = sm(fy);
	fz = sm(fz);
	float v[2][2][2];
	for (int a = 0; a < 2; a++)
		for (int bb = 0; bb < 2; bb++)
			for (int c = 0; c < 2; c++) v[a][bb][c] = latticeValue(ix + a, iy + bb, iz + c, salt);
	auto lerp = [](float a, float c, float t) { return a + t * (c - a); };
	float x00 = lerp(v[0][0][0], v[1][0][0], fx), x10 = lerp(v[0][1][0], v[1][1][0], fx);
	float x01 = lerp(v[0][0][1], v[1][0][1], fx), x11 = lerp(v[0][1][1], v[1][1][1], fx);
	return lerp(lerp(x00, x10, fy), lerp(x01, x11, fy), fz);
}

void valueNoiseMedium(krr_scene &s, const Json &j) {
	int res		  = (int) j.value("res", 256.0);
	int octaves	  = (int) j.value("octaves", 5.0);
	uint32_t seed = (uint32_t) j.value("seed", 11.0);
	float smax	  = (float) j.value("sigma_t_max", 20.0);
	float albedo  = (float) j.value("albedo", 0.8);
	float g		  = (float) j.value("g", 0.3);
	float thresh  = (float) j.value("threshold", 0.0);
	float bounds[2][3] = {{0, 0, 0}, {1, 1, 1}};
	if (j.contains("bounds")) {
		const auto &B = j.at("bounds").items();
		if (B.size() != 2) throw std::runtime_error("bounds must be [[lo],[hi]]");
		for (int k = 0; k < 2; k++)
			for (int a = 0; a < 3; a++) bounds[k][a] = (float) B[k][a].asNumber();
	}
	if (res < 2 || res > 1024 || octaves < 1 || octaves > 12) throw std::runtime_error("value_noise_grid: bad res/octaves");
	auto grid = std::make_unique<std::vector<float>>((size_t) res * res * res);
	float *d  = grid->data();
	unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nt; t++)
		th.emplace_back([=]() {
			for (int z = (int) t; z < res; z += (int) nt)
				for (int y = 0; y < res; y++)
					for (int x = 0; x < res; x++) {
						float sum = 0, amp = 1, norm = 0, freq = 4;
						for (int o = 0; o < octaves; o++) {
							sum += amp * valueNoise((x + 0.5f) / res * freq, (y + 0.5f) / res * freq, (z + 0.5f) / res * freq, seed + 101u * (uint32_t) o);
							norm += amp;
							amp *= 0.5f;
							freq *= 2.f;
						}
						d[(size_t) x + (size_t) res * ((size_t) y + (size_t) res * z)] = std::max(0.f, sum / norm - thresh);
					}
		});
	for (auto &t : th) t.join();
	float mx = 0;
	for (size_t i = 0; i < grid->size(); i++) mx = std::max(mx, d[i]);
	if (mx > 0)
		for (size_t i = 0; i < grid->size(); i++) d[i] /= mx;
	krr_medium m;
	std::memset(&m, 0, sizeof(m));
	m.type = KRR_MEDIUM_GRID;
	for (int a = 0; a < 3; a++) {
		m.sigma_t[a]	= smax;
		m.albedo[a]		= albedo;
		m.res[a]		= res;
		m.bounds_min[a] = bounds[0][a];
		m.bounds_max[a] = bounds[1][a];
	}
	m.g		  = g;
	m.scale	  = 1.f;
	m.density = grid->data();
	s.densities.push_back(std::move(grid));
	s.media.push_back(m);
}

// Dense grid medium given voxel by voxel: {"type": "dense_grid", "res":
// [x, y, z], "density": [x-fastest values], "bounds": [[lo], [hi]],
// "sigma_t": [3], "albedo": [3], "g", "scale"} (defaults: unit cube, sigma_t
// 1, albedo 0.5, g 0, scale 1).
void denseGridMedium(krr_scene &s, const Json &j) {
	if (!j.contains("res") || !j.contains("density")) throw std::runtime_error("dense_grid: needs res and density");
	const auto &R = j.at("res").items();
	if (R.size() != 3) throw std::runtime_error("dense_grid: res must be [x, y, z]");
	int res[3];
	for (int a = 0; a < 3; a++) {
		res[a] = (int) R[a].asNumber();
		if (res[a] < 1 || res[a] > 1024) throw std::runtime_error("dense_grid: res out of range");
	}
	const auto &D = j.at("density").items();
	const size_t n = (size_t) res[0] * res[1] * res[2];
	if (D.size() != n) throw std::runtime_error("dense_grid: density needs res[0]*res[1]*res[2] values");
	auto grid = std::make_unique<std::vector<float>>(n);
	for (size_t i = 0; i < n; i++) {
		(*grid)[i] = (float) D[i].asNumber();
		if (!((*grid)[i] >= 0.f)) throw std::runtime_error("dense_grid: densities must be >= 0");
	}
	krr_medium m;
	std::memset(&m, 0, sizeof(m));
	m.type = KRR_MEDIUM_GRID;
	float st[3] = {1, 1, 1}, al[3] = {0.5f, 0.5f, 0.5f};
	j.vec<3>("sigma_t", st);
	j.vec<3>("albedo", al);
	float bounds[2][3] = {{0, 0, 0}, {1, 1, 1}};
	if (j.contains("bounds")) {
		const auto &B = j.at("bounds").items();
		if (B.size() != 2) throw std::runtime_error("bounds must be [[lo],[hi]]");
		for (int k = 0; k < 2; k++)
			for (int a = 0; a < 3; a++) bounds[k][a] = (float) B[k][a].asNumber();
	}
	for (int a = 0; a < 3; a++) {
		m.sigma_t[a]	= st[a];
		m.albedo[a]		= al[a];
		m.res[a]		= res[a];
		m.bounds_min[a] = bounds[0][a];
		m.bounds_max[a] = bounds[1][a];
		if (!(bounds[1][a] > bounds[0][a])) throw std::runtime_error("dense_grid: empty bounds");
	}
	m.g		  = (float) j.value("g", 0.0);
	m.scale	  = (float) j.value("scale", 1.0);
	m.density = grid->data();
	s.densities.push_back(std::move(grid));
	s.media.push_back(m);
}

// Homogeneous medium with the reference's keys and defaults
// (krrscene.cpp:72-86): sigma_t 1, albedo 0.5, g 0, Le 0; "sigma_a" /
// "sigma_s" (each 0.5 when only the other is given) replace them with
// sigma_t = sigma_a + sigma_s and albedo = sigma_s / sigma_t.
void homogeneousMedium(krr_scene &s, const Json &j) {
	krr_medium m;
	std::memset(&m, 0, sizeof(m));
	m.type = KRR_MEDIUM_HOMOGENEOUS;
	float st[3] = {1, 1, 1}, al[3] = {0.5f, 0.5f, 0.5f}, le[3] = {0, 0, 0};
	j.vec<3>("sigma_t", st);
	j.vec<3>("albedo", al);
	j.vec<3>("Le", le);
	if (j.contains("sigma_a") || j.contains("sigma_s")) {
		float sa[3] = {0.5f, 0.5f, 0.5f}, ss[3] = {0.5f, 0.5f, 0.5f};
		j.vec<3>("sigma_a", sa);
		j.vec<3>("sigma_s", ss);
		for (int a = 0; a < 3; a++) {
			st[a] = sa[a] + ss[a];
			al[a] = ss[a] / st[a];
		}
	}
	for (int a = 0; a < 3; a++) {
		m.sigma_t[a] = st[a];
		m.albedo[a]	 = al[a];
		m.Le[a]		 = le[a];
	}
	m.g		= (float) j.value("g", 0.0);
	m.scale = 1.f;
	s.media.push_back(m);
}

krr_camera lookAtCamera(V eye, V target, V up, int w, int h, const Json &j) {
	krr_camera c;
	std::memset(&c, 0, sizeof(c));
	c.film_size[0]	 = 42.666667f;
	c.film_size[1]	 = 24.0f;
	c.focal_length	 = 21.f;
	c.focal_distance = 10.f;
	c.lens_radius	 = 0.f;
	c.aspect_ratio	 = (w > 0 && h > 0) ? (float) w / (float) h : 1.777777f;
	c.medium		 = -1;
	if (j.isObject()) {
		c.focal_length	 = (float) j.value("focalLength", (double) c.focal_length);
		c.focal_distance = (float) j.value("focalDistance", (double) c.focal_distance);
		c.lens_radius	 = (float) j.value("lensRadius", (double) c.lens_radius);
		c.aspect_ratio	 = (float) j.value("aspectRatio", (double) c.aspect_ratio);
		c.medium		 = (int) j.value("medium", -1.0);
		j.vec<2>("filmSize", c.film_size);
	}
	V f	 = nrmz(sub(target, eye));
	V r	 = nrmz(crs(f, up));
	V u	 = crs(r, f);
	V bk = mul(f, -1.f);
	// columns: camera x -> r, y -> u, z -> -f; translation eye
	float M[12] = {r.x, u.x, bk.x, eye.x, r.y, u.y, bk.y, eye.y, r.z, u.z, bk.z, eye.z};
	std::memcpy(c.transform, M, sizeof(M));
	return c;
}

// ---------------------------------------------------------------- OBJ/MTL

struct ObjMaterial {
	V kd{1, 1, 1}, ks{0, 0, 0}, ke{0, 0, 0}, tf{1, 1, 1};
	float ns = -1.f, ni = -1.f, d = 1.f;
	bool hasTf = false;
	std::string mapKd, mapKs, mapKe; // texture files (map_Kd, map_Ks, map_Ke), full paths
	// the normal map: "norm", else "map_Bump" / "bump", else "disp" (the
	// reference takes assimp's NORMALS, HEIGHT and DISPLACEMENT slots, in that
	// order, as its Normal texture: assimp.cpp:48-58,66-90)
	std::string mapNorm, mapBump, mapDisp;
};

std::string joinPath(const std::string &dir, const std::string &p) {
	if (p.empty() || p[0] == '/' || dir.empty()) return p;
	return dir.back() == '/' ? dir + p : dir + "/" + p;
}
std::string dirOf(const std::string &p) {
	const size_t k = p.find_last_of('/');
	return k == std::string::npos ? std::string() : p.substr(0, k);
}

void readMtl(const std::string &path, std::map<std::string, ObjMaterial> &out) {
	std::ifstream f(path);
	if (!f) throw std::runtime_error("cannot open material library " + path);
	std::string line, cur;
	auto rd3 = [](std::istringstream &ss) {
		V v{0, 0, 0};
		ss >> v.x >> v.y >> v.z;
		return v;
	};
	while (std::getline(f, line)) {
		std::istringstream ss(line);
		std::string k;
		if (!(ss >> k) || k[0] == '#') continue;
		if (k == "newmtl") {
			ss >> cur;
			out[cur] = ObjMaterial{};
			continue;
		}
		if (cur.empty()) continue;
		ObjMaterial &m = out[cur];
		if (k == "Kd") m.kd = rd3(ss);
		else if (k == "Ks") m.ks = rd3(ss);
		else if (k == "Ke") m.ke = rd3(ss);
		else if (k == "Tf") {
			m.tf	= rd3(ss);
			m.hasTf = true;
		} else if (k == "Ns") ss >> m.ns;
		else if (k == "Ni") ss >> m.ni;
		else if (k == "d") ss >> m.d;
		else if (k == "Tr") {
			float tr = 0.f;
			ss >> tr;
			m.d = 1.f - tr;
		} else if (k == "map_Kd" || k == "map_Ks" || k == "map_Ke") { // options first, the file name last
			std::string tok, file;
			while (ss >> tok) file = tok;
			if (!file.empty()) (k == "map_Kd" ? m.mapKd : k == "map_Ks" ? m.mapKs : m.mapKe) = joinPath(dirOf(path), file);
		} else if (k == "norm" || k == "map_Bump" || k == "map_bump" || k == "bump" || k == "disp") {
			std::string tok, file;
			while (ss >> tok) file = tok;
			if (!file.empty()) (k == "norm" ? m.mapNorm : k == "disp" ? m.mapDisp : m.mapBump) = joinPath(dirOf(path), file);
		}
	}
}

void readImage(const std::string &path, std::vector<float> &rgba, int &w, int &h);

// 1 + the index of the texture read from this image file (each file once).
int32_t textureFor(krr_scene &s, const std::string &path) {
	auto it = s.textureByPath.find(path);
	if (it != s.textureByPath.end()) return it->second;
	auto img = std::make_unique<std::vector<float>>();
	krr_texture t;
	readImage(path, *img, t.width, t.height);
	t.rgba = img->data();
	s.images.push_back(std::move(img));
	s.textures.push_back(t);
	const int32_t id	 = (int32_t) s.textures.size();
	s.textureByPath[path] = id;
	return id;
}

// AssimpImporter's OBJ material (assimp.cpp:60-65,96-224): shininess Ns ->
// roughness sqrt(2 / (Ns + 2)) (0 from 1000 on), specular.a = 1 - roughness;
// Ni -> IoR; opacity 0 or a fully transparent Tf -> dielectric with
// specular transmission; Kd, Ks; Ke -> emission.  The reference's default BSDF
// for file materials is Disney, which this build does not have: such
// materials become Lambertian with the same diffuse colour.
krr_material objMaterial(const ObjMaterial &o) {
	krr_material m = makeMaterial(KRR_MAT_DIFFUSE, o.kd, 1.f, 1.5f, 0.f);
	m.specular[0]  = o.ks.x;
	m.specular[1]  = o.ks.y;
	m.specular[2]  = o.ks.z;
	m.specular[3]  = 0.f;
	if (o.ns >= 0.f) {
		const float r = o.ns >= 1000.f ? 0.f : std::min(1.f, std::max(0.f, std::sqrt(2.f / (o.ns + 2.f))));
		m.specular[3] = 1.f - r;
	}
	if (o.ni >= 0.f) m.ior = o.ni;
	if (o.d == 0.f) {
		m.specular_transmission = 1.f;
		m.type					= KRR_MAT_DIELECTRIC;
	}
	if (o.hasTf) {
		const float t = 0.2126f * (1.f - o.tf.x) + 0.7152f * (1.f - o.tf.y) + 0.0722f * (1.f - o.tf.z); // luminance of 1 - Tf
		m.specular_transmission = t;
		if (t > 1.f - 1e-6f) m.type = KRR_MAT_DIELECTRIC;
	}
	return m;
}

// Node transform (krrscene.cpp:317-333, scene graph nodes): translate,
// rotate (quaternion [w, x, y, z]), scale, p' = T + R (S p) for one node;
// nested nodes compose parent * child as the scene graph does.  Stored as
// the affine map p' = L p + t.
struct NodeXform {
	float L[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
	V t{0, 0, 0};
	bool identity() const {
		static const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
		return std::memcmp(L, I, sizeof(I)) == 0 && t.x == 0.f && t.y == 0.f && t.z == 0.f;
	}
	V linear(V p) const { return {L[0] * p.x + L[1] * p.y + L[2] * p.z, L[3] * p.x + L[4] * p.y + L[5] * p.z, L[6] * p.x + L[7] * p.y + L[8] * p.z}; }
	V point(V p) const { return add(linear(p), t); }
	V normal(V n) const { // inverse transpose of L: its cofactor matrix over det(L)
		const float c[9] = {L[4] * L[8] - L[5] * L[7], L[5] * L[6] - L[3] * L[8], L[3] * L[7] - L[4] * L[6],
							L[2] * L[7] - L[1] * L[8], L[0] * L[8] - L[2] * L[6], L[1] * L[6] - L[0] * L[7],
							L[1] * L[5] - L[2] * L[4], L[2] * L[3] - L[0] * L[5], L[0] * L[4] - L[1] * L[3]};
		const float det = L[0] * c[0] + L[1] * c[1] + L[2] * c[2];
		const float sg	= det < 0.f ? -1.f : 1.f;
		return nrmz({sg * (c[0] * n.x + c[1] * n.y + c[2] * n.z), sg * (c[3] * n.x + c[4] * n.y + c[5] * n.z), sg * (c[6] * n.x + c[7] * n.y + c[8] * n.z)});
	}
	NodeXform then(const NodeXform &child) const { // this * child
		NodeXform r;
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++) r.L[3 * i + j] = L[3 * i] * child.L[j] + L[3 * i + 1] * child.L[3 + j] + L[3 * i + 2] * child.L[6 + j];
		r.t = point(child.t);
		return r;
	}
};

NodeXform nodeXform(const Json &m) {
	NodeXform x;
	float tr[3] = {0, 0, 0}, sc[3] = {1, 1, 1}, q[4] = {1, 0, 0, 0};
	m.vec<3>("translate", tr);
	m.vec<3>("scale", sc);
	m.vec<4>("rotate", q);
	const float w = q[0], a = q[1], b = q[2], c = q[3];
	const float n = std::sqrt(w * w + a * a + b * b + c * c);
	if (!(n > 0.f)) throw std::runtime_error("rotate: the quaternion has zero length");
	const float W = w / n, X = a / n, Y = b / n, Z = c / n;
	const float R[9] = {1 - 2 * (Y * Y + Z * Z), 2 * (X * Y - Z * W),	  2 * (X * Z + Y * W),
						2 * (X * Y + Z * W),	 1 - 2 * (X * X + Z * Z), 2 * (Y * Z - X * W),
						2 * (X * Z - Y * W),	 2 * (Y * Z + X * W),	  1 - 2 * (X * X + Y * Y)};
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) x.L[3 * i + j] = R[3 * i + j] * sc[j];
	x.t = {tr[0], tr[1], tr[2]};
	return x;
}

// Per-vertex tangents of a mesh with texture coordinates, standing for
// assimp's aiProcess_CalcTangentSpace (assimp.cpp:249) as the reference
// stores them (assimp.cpp:376-383: made orthogonal to the vertex normal and
// normalised).  Per triangle the tangent is the direction of increasing u
// (Lengyel's construction from the edge vectors and texture-coordinate
// deltas); each vertex sums the tangents of its triangles.  A vertex whose
// triangles have no usable texture-coordinate gradient gets a unit vector
// perpendicular to n (n crossed with the axis of n's smallest component).
std::vector<V> uvTangents(const std::vector<V> &verts, const std::vector<V> &nrms, const std::vector<float> &uvs, const std::vector<int> &tris) {
	std::vector<V> acc(verts.size(), V{0, 0, 0});
	for (size_t t = 0; t + 2 < tris.size(); t += 3) {
		const int a = tris[t], b = tris[t + 1], c = tris[t + 2];
		const V e1 = sub(verts[b], verts[a]), e2 = sub(verts[c], verts[a]);
		const float du1 = uvs[2 * b] - uvs[2 * a], dv1 = uvs[2 * b + 1] - uvs[2 * a + 1];
		const float du2 = uvs[2 * c] - uvs[2 * a], dv2 = uvs[2 * c + 1] - uvs[2 * a + 1];
		const float det = du1 * dv2 - du2 * dv1;
		if (det == 0.f || !std::isfinite(det)) continue;
		const V tg = mul(sub(mul(e1, dv2), mul(e2, dv1)), 1.f / det);
		for (int v : {a, b, c}) acc[v] = add(acc[v], tg);
	}
	std::vector<V> out(verts.size());
	for (size_t i = 0; i < verts.size(); i++) {
		const V n = nrms[i];
		V t		  = sub(acc[i], mul(n, dt(n, acc[i])));
		if (dt(t, t) <= 1e-20f * std::max(dt(acc[i], acc[i]), 1e-30f) || !std::isfinite(dt(t, t))) {
			const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
			const V e = ax <= ay && ax <= az ? V{1, 0, 0} : ay <= az ? V{0, 1, 0} : V{0, 0, 1};
			t		  = crs(n, e);
		}
		out[i] = nrmz(t);
	}
	return out;
}

// Wavefront OBJ: v, vn, f (v, v/t, v//n, v/t/n, negative indices, polygons
// as fans), usemtl / mtllib, o / g.  One mesh per run of faces with one
// material and one object or group name; the mesh takes that name (the name
// a medium's "meshes" list refers to).
// Meshes without vn get smooth normals as assimp's GenSmoothNormals
// (assimp.cpp:248-261; it runs before vertices are joined, on one vertex per
// face corner): per position, the normalised sum of the unit normals of the
// mesh's faces (OBJ polygons) that touch it.  "materials" entries override
// file materials by name.
void loadObj(Builder &b, const std::string &path, const NodeXform &xf, const std::map<std::string, krr_material> &overrides) {
	std::ifstream f(path);
	if (!f) throw std::runtime_error("cannot open model " + path);
	std::map<std::string, ObjMaterial> mtl;
	std::map<std::string, int> matIndex;
	std::vector<V> P, N;
	std::vector<float> T; // vt: u, 1 - v (aiProcess_FlipUVs, assimp.cpp:264)
	struct Corner {
		int p, n, t;
	};
	struct Run {
		std::string mat, name;
		std::vector<Corner> tri;  // 3 per triangle
		std::map<int, V> smooth; // position -> summed unit polygon normals
	};
	std::vector<Run> runs(1);
	std::string line;
	auto resolve = [](int i, size_t n) { return i > 0 ? i - 1 : (int) n + i; };
	while (std::getline(f, line)) {
		std::istringstream ss(line);
		std::string k;
		if (!(ss >> k) || k[0] == '#') continue;
		if (k == "v") {
			V v{0, 0, 0};
			ss >> v.x >> v.y >> v.z;
			P.push_back(v);
		} else if (k == "vn") {
			V v{0, 0, 0};
			ss >> v.x >> v.y >> v.z;
			N.push_back(v);
		} else if (k == "vt") {
			float u = 0.f, v = 0.f;
			ss >> u >> v;
			T.push_back(u);
			T.push_back(1.f - v);
		} else if (k == "mtllib") {
			std::string lib;
			std::getline(ss, lib);
			lib.erase(0, lib.find_first_not_of(" \t"));
			lib.erase(lib.find_last_not_of(" \t\r") + 1);
			readMtl(joinPath(dirOf(path), lib), mtl);
		} else if (k == "usemtl") {
			std::string name;
			ss >> name;
			if (runs.back().tri.empty()) runs.back().mat = name;
			else runs.push_back(Run{name, runs.back().name, {}, {}});
		} else if (k == "o" || k == "g") {
			std::string name;
			std::getline(ss, name);
			name.erase(0, name.find_first_not_of(" \t"));
			name.erase(name.find_last_not_of(" \t\r") + 1);
			if (runs.back().tri.empty()) runs.back().name = name;
			else runs.push_back(Run{runs.back().mat, name, {}, {}});
		} else if (k == "f") {
			std::vector<Corner> poly;
			std::string tok;
			while (ss >> tok) {
				int vi = 0, ni = 0, ti = 0;
				const size_t s1 = tok.find('/');
				vi				= std::atoi(tok.substr(0, s1).c_str());
				if (s1 != std::string::npos) {
					const size_t s2 = tok.find('/', s1 + 1);
					ti				= std::atoi(tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1).c_str());
					if (s2 != std::string::npos && s2 + 1 < tok.size()) ni = std::atoi(tok.substr(s2 + 1).c_str());
				}
				Corner c{resolve(vi, P.size()), ni ? resolve(ni, N.size()) : -1, ti ? resolve(ti, T.size() / 2) : -1};
				if (c.p < 0 || c.p >= (int) P.size() || c.n >= (int) N.size() || 2 * c.t >= (int) T.size() || (ti && c.t < 0) || (ni && c.n < 0))
					throw std::runtime_error(path + ": face index out of range");
				poly.push_back(c);
			}
			if (poly.size() < 3) continue;
			for (size_t i = 2; i < poly.size(); i++) { // fan
				runs.back().tri.push_back(poly[0]);
				runs.back().tri.push_back(poly[i - 1]);
				runs.back().tri.push_back(poly[i]);
			}
			// the polygon's normal (of its first triangle), once per position
			const V p0 = xf.point(P[poly[0].p]), p1 = xf.point(P[poly[1].p]), p2 = xf.point(P[poly[2].p]);
			const V fn = nrmz(crs(sub(p1, p0), sub(p2, p0)));
			for (size_t k = 0; k < poly.size(); k++) {
				bool dup = false;
				for (size_t q = 0; q < k; q++) dup |= poly[q].p == poly[k].p;
				if (!dup) runs.back().smooth[poly[k].p] = add(runs.back().smooth[poly[k].p], fn);
			}
		}
	}
	for (const Run &r : runs) {
		if (r.tri.empty()) continue;
		// material: JSON override by name, else the MTL entry, else default
		int mi;
		auto mit = matIndex.find(r.mat);
		if (mit != matIndex.end()) mi = mit->second;
		else {
			auto ov = overrides.find(r.mat);
			const auto ot = mtl.find(r.mat);
			const ObjMaterial om = ot != mtl.end() ? ot->second : ObjMaterial{};
			krr_material fm	  = objMaterial(om);
			fm.diffuse_texture	= om.mapKd.empty() ? 0 : textureFor(b.s, om.mapKd);
			fm.specular_texture = om.mapKs.empty() ? 0 : textureFor(b.s, om.mapKs);
			fm.emissive_texture = om.mapKe.empty() ? 0 : textureFor(b.s, om.mapKe);
			const std::string &nm = !om.mapNorm.empty() ? om.mapNorm : !om.mapBump.empty() ? om.mapBump : om.mapDisp;
			fm.normal_texture	  = nm.empty() ? 0 : textureFor(b.s, nm);
			krr_material mm = fm;
			if (ov != overrides.end()) { // the JSON entry sets the params; the file's textures stay (krrscene.cpp:199-205)
				mm = ov->second;
				if (!mm.diffuse_texture) mm.diffuse_texture = fm.diffuse_texture;
				if (!mm.specular_texture) mm.specular_texture = fm.specular_texture;
				if (!mm.emissive_texture) mm.emissive_texture = fm.emissive_texture;
				if (!mm.normal_texture) mm.normal_texture = fm.normal_texture;
			}
			mi				= b.addMaterial(mm);
			matIndex[r.mat] = mi;
		}
		const auto ot	 = mtl.find(r.mat);
		// Ke becomes a constant Emissive texture (assimp.cpp:168-180); the mesh
		// light takes scale = max(Ke) and evaluates that texture, so it emits
		// scale * Ke, or scale * map_Ke(uv) when the file has one
		// (mesh.cpp:39-59, light.h:188-193)
		V emission = ot != mtl.end() ? ot->second.ke : V{0, 0, 0};
		if (ot != mtl.end() && ot->second.mapKe.empty()) emission = mul(emission, std::max(std::max(emission.x, emission.y), emission.z));
		// vertices of the run: one per (position, normal) corner pair
		std::map<std::tuple<int, int, int>, int> vid;
		std::vector<V> verts, nrms;
		std::vector<float> uvs;
		std::vector<int> tris;
		bool haveN = true, haveT = true;
		for (const Corner &c : r.tri) {
			haveN &= c.n >= 0;
			haveT &= c.t >= 0;
		}
		const std::map<int, V> &smooth = r.smooth;
		for (const Corner &c : r.tri) {
			const auto key = std::make_tuple(c.p, haveN ? c.n : -1, haveT ? c.t : -1);
			auto it		   = vid.find(key);
			if (it == vid.end()) {
				it = vid.emplace(key, (int) verts.size()).first;
				verts.push_back(xf.point(P[c.p]));
				nrms.push_back(haveN ? xf.normal(N[c.n]) : nrmz(smooth.at(c.p)));
				if (haveT) {
					uvs.push_back(T[2 * c.t]);
					uvs.push_back(T[2 * c.t + 1]);
				}
			}
			tris.push_back(it->second);
		}
		b.addMesh(verts, nrms, tris, mi, emission, -1, -1, r.name, uvs, haveT ? uvTangents(verts, nrms, uvs, tris) : std::vector<V>{});
	}
}

// A "materials" entry (krrscene.cpp:198-229).
krr_material jsonMaterial(const Json &m) {
	const Json &p	 = m.contains("params") ? m.at("params") : m;
	std::string bsdf = m.value("bsdf", "diffuse");
	int type		 = bsdf == "diffuse" ? KRR_MAT_DIFFUSE : bsdf == "dielectric" ? KRR_MAT_DIELECTRIC
									   : bsdf == "conductor" ? KRR_MAT_CONDUCTOR : bsdf == "null" ? KRR_MAT_NULL : -1;
	if (type < 0) throw std::runtime_error("unsupported bsdf \"" + bsdf + "\"");
	krr_material mm;
	std::memset(&mm, 0, sizeof(mm));
	mm.type		 = type;
	float dif[3] = {1, 1, 1}, spec[3] = {0, 0, 0};
	p.vec<3>("diffuse", dif);
	p.vec<3>("specular", spec);
	for (int a = 0; a < 3; a++) {
		mm.diffuse[a]  = dif[a];
		mm.specular[a] = spec[a];
	}
	mm.specular[3]			 = 1.f - (float) p.value("roughness", 1.0);
	mm.specular_transmission = (float) p.value("specular_transmission", 0.0);
	mm.anisotropic			 = (float) p.value("anisotropic", 0.0);
	mm.ior					 = (float) p.value("eta", 1.5);
	return mm;
}

// Camera of KiRaRay's scene JSON: CameraData keys under "mData"
// (camera.h:64-81) and the orbit controller (camera.cpp:48-61): rotation
// Ry(yaw) Rx(pitch), forward = R (0, 0, -1), position = target - forward *
// radius.  Without a controller the camera sits at the origin looking down -z.
krr_camera referenceCamera(const Json &cam, const Json *ctrl) {
	krr_camera c;
	std::memset(&c, 0, sizeof(c));
	c.film_size[0]	 = 42.666667f;
	c.film_size[1]	 = 24.0f;
	c.focal_length	 = 21.f;
	c.focal_distance = 10.f;
	c.lens_radius	 = 0.f;
	c.aspect_ratio	 = 1.777777f;
	c.medium		 = -1;
	const Json d	 = cam.contains("mData") ? cam.at("mData") : Json::object();
	c.focal_length	 = (float) d.value("focalLength", (double) c.focal_length);
	c.focal_distance = (float) d.value("focalDistance", (double) c.focal_distance);
	c.lens_radius	 = (float) d.value("lensRadius", (double) c.lens_radius);
	c.aspect_ratio	 = (float) d.value("aspectRatio", (double) c.aspect_ratio);
	c.shutter_open	 = (float) d.value("shutterOpen", 0.0);
	c.shutter_time	 = (float) d.value("shutterTime", 0.0);
	float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, pos[3] = {0, 0, 0};
	if (ctrl) {
		const Json cd	 = ctrl->contains("mData") ? ctrl->at("mData") : Json::object();
		float target[3] = {0, 0, 0};
		cd.vec<3>("target", target);
		const float radius = (float) cd.value("radius", 5.0), pitch = (float) cd.value("pitch", 0.0), yaw = (float) cd.value("yaw", 0.0);
		const float cy = std::cos(yaw), sy = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch);
		// Ry(yaw) * Rx(pitch)
		const float M[9] = {cy, sy * sp, sy * cp, 0, cp, -sp, -sy, cy * sp, cy * cp};
		std::memcpy(R, M, sizeof(M));
		const V fwd = {-R[2], -R[5], -R[8]};
		pos[0]		= target[0] - fwd.x * radius;
		pos[1]		= target[1] - fwd.y * radius;
		pos[2]		= target[2] - fwd.z * radius;
	}
	const float T[12] = {R[0], R[1], R[2], pos[0], R[3], R[4], R[5], pos[1], R[6], R[7], R[8], pos[2]};
	std::memcpy(c.transform, T, sizeof(T));
	return c;
}

// Float images for infinite lights, as RGBA rows from the top of the image
// (the reference loads images without a vertical flip, texture.cpp:27-60).
// Portable float maps (.pfm, "PF" RGB or "Pf" grey, rows stored bottom-up)
// and NumPy arrays (.npy, little-endian float32, H x W, H x W x 3 or
// H x W x 4, C order).  The reference also reads .exr / .hdr / .png through
// image libraries this build does not link.
void readImage(const std::string &path, std::vector<float> &rgba, int &w, int &h) {
	std::ifstream f(path, std::ios::binary);
	if (!f) throw std::runtime_error("cannot open image " + path);
	const std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
	auto fail = [&](const char *why) { throw std::runtime_error("image " + path + ": " + why); };
	int ch = 0;
	const char *px = nullptr;
	bool bottomUp = false, bigEndian = false;
	if (data.size() >= 2 && data[0] == 'P' && (data[1] == 'F' || data[1] == 'f')) {
		ch = data[1] == 'F' ? 3 : 1;
		std::istringstream hs(data);
		std::string magic;
		double scale = 0;
		hs >> magic >> w >> h >> scale;
		if (!hs || w < 1 || h < 1) fail("bad PFM header");
		const size_t off = (size_t) hs.tellg() + 1; // one whitespace byte after the scale
		if (off + (size_t) w * h * ch * 4 > data.size()) fail("truncated PFM data");
		px		  = data.data() + off;
		bottomUp  = true;
		bigEndian = scale > 0;
	} else if (data.size() >= 10 && (unsigned char) data[0] == 0x93 && data.compare(1, 5, "NUMPY") == 0) {
		const size_t hl	 = data[6] == 1 ? (size_t) (unsigned char) data[8] | (size_t) (unsigned char) data[9] << 8
										: (size_t) (unsigned char) data[8] | (size_t) (unsigned char) data[9] << 8 | (size_t) (unsigned char) data[10] << 16;
		const size_t off = data[6] == 1 ? 10 + hl : 12 + hl;
		const std::string hdr = data.substr(data[6] == 1 ? 10 : 12, hl);
		if (hdr.find("'<f4'") == std::string::npos) fail("NPY data must be little-endian float32");
		if (hdr.find("'fortran_order': False") == std::string::npos) fail("NPY data must be in C order");
		const size_t a = hdr.find('('), z = hdr.find(')');
		if (a == std::string::npos || z == std::string::npos) fail("bad NPY shape");
		std::vector<long> dims;
		std::istringstream ds(hdr.substr(a + 1, z - a - 1));
		std::string tok;
		while (std::getline(ds, tok, ','))
			if (tok.find_first_not_of(" ") != std::string::npos) dims.push_back(std::stol(tok));
		if (dims.size() == 2) ch = 1;
		else if (dims.size() == 3 && (dims[2] == 3 || dims[2] == 4)) ch = (int) dims[2];
		else fail("NPY shape must be (H, W), (H, W, 3) or (H, W, 4)");
		h = (int) dims[0];
		w = (int) dims[1];
		if (w < 1 || h < 1 || off + (size_t) w * h * ch * 4 > data.size()) fail("truncated NPY data");
		px = data.data() + off;
	} else fail("only .pfm and .npy float images are read");
	if (w > 65536 || h > 65536) fail("larger than 65536 texels on a side");
	rgba.assign((size_t) w * h * 4, 1.f);
	for (int y = 0; y < h; y++) {
		const int sy = bottomUp ? h - 1 - y : y;
		for (int x = 0; x < w; x++)
			for (int c = 0; c < 4; c++) {
				const int sc = ch == 1 ? 0 : c;
				if (c >= ch && !(ch == 1 && c < 3)) continue;
				unsigned char b[4];
				std::memcpy(b, px + 4 * (((size_t) sy * w + x) * ch + sc), 4);
				if (bigEndian) std::swap(b[0], b[3]), std::swap(b[1], b[2]);
				float v;
				std::memcpy(&v, b, 4);
				rgba[4 * ((size_t) y * w + x) + c] = v;
			}
	}
}

// An infinite light (light.h:208-259): tint * scale, with an image file and
// the rotation R (row-major, world from light) when given.
void addInfinite(krr_scene &s, const float tint[3], float scale, const std::string &image, const float *R) {
	krr_infinite_light il;
	std::memset(&il, 0, sizeof(il));
	for (int a = 0; a < 3; a++) il.tint[a] = tint[a];
	il.scale = scale;
	if (R)
		for (int k = 0; k < 9; k++) il.rotation[k] = R[k];
	if (!image.empty()) {
		auto img = std::make_unique<std::vector<float>>();
		readImage(image, *img, il.image_width, il.image_height);
		il.image_rgba = img->data();
		s.images.push_back(std::move(img));
	}
	s.infinite.push_back(il);
}

// ---------------------------------------------------------------------------
// OpenVDB volumes.
//
// Reader of float grids in the .vdb files OpenVDB's io::File writes (file
// format 222 and later), standing for openvdb::io::File::readGrid followed by
// nanovdb::openToNanoVDB in the reference (util/volume.cpp:46-78).  A grid is
// read in the order io::File writes it: its compression flags, metadata and
// transform, the 5-4-3 tree topology (root tiles and children; internal nodes
// with child mask, value mask and tile values; leaf value masks), then the
// leaf buffers.  Read are Tree_float_5_4_3 grids (values saved as float or
// half) with a (uniform) ScaleMap or ScaleTranslateMap and no compression or
// active-mask compression; zip / blosc compressed grids, other value types and
// transforms are errors.  The grid becomes a dense grid over the bounding box
// of its active voxels and active tiles: voxel (i, j, k) of that box holds the
// tree's value at index lo + (i, j, k), which is what the reference's
// trilinear sampler reads there (inactive voxels keep their stored values;
// outside the box every value is the zero background).
// ---------------------------------------------------------------------------
struct VdbGrid {
	int lo[3], hi[3];			// index box of the active voxels, inclusive
	double voxel[3], origin[3]; // index -> world (before node transforms): origin + voxel * index
	std::vector<float> dense;	// (hi - lo + 1) voxels per axis, x fastest
};

struct VdbIn {
	const std::vector<uint8_t> &d;
	size_t pos;
	const std::string &path;
	void need(size_t n) const {
		if (n > d.size() || pos > d.size() - n) throw std::runtime_error(path + ": truncated OpenVDB file");
	}
	template <class T> T get() {
		need(sizeof(T));
		T v;
		std::memcpy(&v, d.data() + pos, sizeof(T));
		pos += sizeof(T);
		return v;
	}
	std::string str() {
		const uint32_t n = get<uint32_t>();
		need(n);
		std::string s(reinterpret_cast<const char *>(d.data() + pos), n);
		pos += n;
		return s;
	}
	void skip(size_t n) {
		need(n);
		pos += n;
	}
	// MetaMap: count, then per entry the name, the type name and the value's bytes
	void skipMeta() {
		const uint32_t n = get<uint32_t>();
		for (uint32_t i = 0; i < n; i++) {
			str();
			str();
			skip(get<uint32_t>());
		}
	}
	// NodeMask of `bits` bits stored as 64-bit words, bit i in word i / 64 at bit i % 64
	std::vector<uint8_t> mask(int bits) {
		need((size_t) bits / 8);
		std::vector<uint8_t> m((size_t) bits);
		for (int i = 0; i < bits; i++) m[(size_t) i] = (d[pos + (size_t) (i >> 3)] >> (i & 7)) & 1u;
		pos += (size_t) bits / 8;
		return m;
	}
};

// IEEE binary16 -> binary32 (exact).
float halfToFloat(uint16_t h) {
	const uint32_t s = (uint32_t) (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, f = h & 0x3ffu;
	uint32_t u;
	if (e == 0 && f == 0) u = s;
	else if (e == 0) { // subnormal: shift the fraction up to a leading 1
		uint32_t m = f, ex = 127 - 15 + 1;
		while (!(m & 0x400u)) {
			m <<= 1;
			ex--;
		}
		u = s | (ex << 23) | ((m & 0x3ffu) << 13);
	} else if (e == 31) u = s | 0x7f800000u | (f << 13);
	else u = s | ((e + 127 - 15) << 23) | (f << 13);
	float r;
	std::memcpy(&r, &u, 4);
	return r;
}

// One node's values as io::readCompressedValues stores them: a metadata byte
// naming how inactive values are kept (0: background, 1: -background, 2: one
// stored value, 3-5: a selection mask picking between two of them, 6: all
// values stored), the stored inactive values (full floats), the selection
// mask, then the active values only (with active-mask compression) or all n,
// as half or float.
std::vector<float> vdbValues(VdbIn &in, int n, const std::vector<uint8_t> &valueMask, bool maskCompressed, bool half, float bg) {
	const int8_t meta = in.get<int8_t>();
	if (meta < 0 || meta > 6) throw std::runtime_error(in.path + ": bad OpenVDB value metadata");
	float inactive0 = meta == 0 ? bg : -bg, inactive1 = bg;
	if (meta == 2 || meta == 4 || meta == 5) {
		inactive0 = in.get<float>();
		if (meta == 5) inactive1 = in.get<float>();
	}
	std::vector<uint8_t> sel;
	if (meta == 3 || meta == 4 || meta == 5) sel = in.mask(n);
	int count = n;
	if (maskCompressed && meta != 6) {
		count = 0;
		for (int i = 0; i < n; i++) count += valueMask[(size_t) i];
	}
	std::vector<float> vals((size_t) count);
	for (int i = 0; i < count; i++) vals[(size_t) i] = half ? halfToFloat(in.get<uint16_t>()) : in.get<float>();
	if (count == n) return vals;
	std::vector<float> out((size_t) n);
	for (int i = 0, k = 0; i < n; i++)
		out[(size_t) i] = valueMask[(size_t) i] ? vals[(size_t) k++] : (!sel.empty() && sel[(size_t) i] ? inactive1 : inactive0);
	return out;
}

struct VdbFile {
	struct Desc {
		std::string name, type, instanceParent;
		int64_t gridPos, blockPos, endPos;
	};
	std::vector<uint8_t> bytes;
	std::vector<Desc> grids;
	std::string path;
};

// File header and grid descriptors (io::File::open).
VdbFile openVdb(const std::string &path) {
	VdbFile f;
	f.path = path;
	std::ifstream is(path, std::ios::binary);
	if (!is) throw std::runtime_error("cannot open OpenVDB file " + path);
	f.bytes.assign(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
	VdbIn in{f.bytes, 0, f.path};
	if (in.get<int64_t>() != 0x56444220) throw std::runtime_error(path + ": not an OpenVDB file");
	if (in.get<uint32_t>() < 222) throw std::runtime_error(path + ": OpenVDB file format older than 222 is not read");
	in.get<uint32_t>(); // library major, minor
	in.get<uint32_t>();
	if (!in.get<uint8_t>()) throw std::runtime_error(path + ": OpenVDB file without grid offsets is not read");
	in.skip(36); // UUID
	in.skipMeta();
	// Each grid's descriptor sits in front of the grid's data; the next
	// descriptor follows the grid's end.
	const int32_t nGrids = in.get<int32_t>();
	for (int32_t g = 0; g < nGrids; g++) {
		VdbFile::Desc d;
		d.name = in.str();
		d.name = d.name.substr(0, d.name.find('\x1e')); // unique-name suffix of repeated names
		d.type = in.str();
		d.instanceParent = in.str();
		d.gridPos		 = in.get<int64_t>();
		d.blockPos		 = in.get<int64_t>();
		d.endPos		 = in.get<int64_t>();
		if (d.gridPos < (int64_t) in.pos || d.blockPos < d.gridPos || d.endPos < d.blockPos || (size_t) d.endPos > f.bytes.size())
			throw std::runtime_error(path + ": bad grid offsets");
		f.grids.push_back(d);
		in.pos = (size_t) d.endPos;
	}
	return f;
}

bool vdbHasGrid(const VdbFile &f, const std::string &name) {
	for (const auto &d : f.grids)
		if (d.name == name) return true;
	return false;
}

VdbGrid readVdbGrid(const VdbFile &f, const std::string &key) {
	const VdbFile::Desc *desc = nullptr;
	for (const auto &d : f.grids)
		if (d.name == key) {
			desc = &d;
			break;
		}
	const std::string &path = f.path;
	if (!desc) throw std::runtime_error(path + ": no grid named \"" + key + "\"");
	const bool half = desc->type == "Tree_float_5_4_3_HalfFloat";
	if (!half && desc->type != "Tree_float_5_4_3")
		throw std::runtime_error(path + ": grid \"" + key + "\" of type " + desc->type + " is not read (float grids only)");
	if (!desc->instanceParent.empty()) throw std::runtime_error(path + ": instanced grid \"" + key + "\" is not read");
	VdbIn in{f.bytes, (size_t) desc->gridPos, path};
	const uint32_t compression = in.get<uint32_t>();
	if (compression & ~2u) throw std::runtime_error(path + ": zip or blosc compressed grids are not read");
	const bool maskCompressed = (compression & 2u) != 0;
	in.skipMeta();
	VdbGrid g;
	const std::string map = in.str();
	if (map == "UniformScaleTranslateMap" || map == "ScaleTranslateMap") {
		// translation, scale, voxel size, then three derived vectors
		double rec[18];
		for (double &x : rec) x = in.get<double>();
		for (int a = 0; a < 3; a++) {
			g.origin[a] = rec[a];
			g.voxel[a]	= rec[3 + a];
		}
	} else if (map == "UniformScaleMap" || map == "ScaleMap") {
		// scale, voxel size, then three derived vectors
		double rec[15];
		for (double &x : rec) x = in.get<double>();
		for (int a = 0; a < 3; a++) {
			g.origin[a] = 0.0;
			g.voxel[a]	= rec[a];
		}
	} else throw std::runtime_error(path + ": transform " + map + " is not read (scale and scale-translate maps only)");
	for (int a = 0; a < 3; a++)
		if (!(g.voxel[a] > 0.0)) throw std::runtime_error(path + ": transforms with non-positive scales are not read");
	if (in.get<int32_t>() != 1) throw std::runtime_error(path + ": trees with more than one buffer are not read");
	const float bg = in.get<float>();
	if (bg != 0.f) throw std::runtime_error(path + ": grids with a non-zero background are not read");
	const uint32_t nTiles = in.get<uint32_t>(), nChildren = in.get<uint32_t>();
	struct Block { // a leaf (span 8) or a tile (span 8, 128 or 4096)
		int o[3], span;
		float v;
		bool active;
	};
	std::vector<Block> leaves, tiles;
	// node origins are added to spans of up to 4096 voxels below: far ones
	// would overflow int (no real grid comes near 2^28 voxels from the origin)
	auto origin = [&](int *o) {
		for (int a = 0; a < 3; a++) {
			o[a] = in.get<int32_t>();
			if (o[a] < -(1 << 28) || o[a] > (1 << 28)) throw std::runtime_error(path + ": node origin out of range");
		}
	};
	for (uint32_t t = 0; t < nTiles; t++) {
		Block b;
		origin(b.o);
		b.v		 = in.get<float>();
		b.active = in.get<uint8_t>() != 0;
		b.span	 = 4096;
		if (b.v != 0.f) throw std::runtime_error(path + ": root tiles with non-zero values are not read");
		if (b.active) tiles.push_back(b);
	}
	// internal node of 2^(3 L) entries, each covering 2^(3 C) voxels (L = 5, C = 7 and L = 4, C = 3)
	auto internal = [&](auto &&self, int L, const int *o) -> void {
		const int n = 1 << (3 * L), C = L == 5 ? 7 : 3, m = (1 << L) - 1;
		const std::vector<uint8_t> childMask = in.mask(n), valueMask = in.mask(n);
		const std::vector<float> vals		 = vdbValues(in, n, valueMask, maskCompressed, half, bg);
		for (int i = 0; i < n; i++) {
			const int co[3] = {o[0] + ((i >> (2 * L)) << C), o[1] + (((i >> L) & m) << C), o[2] + ((i & m) << C)};
			if (childMask[(size_t) i]) {
				if (L == 5) self(self, 4, co);
				else {
					in.skip(64); // the leaf's value mask, read again with its buffer
					leaves.push_back(Block{{co[0], co[1], co[2]}, 8, 0.f, true});
				}
			} else if (valueMask[(size_t) i] || vals[(size_t) i] != 0.f)
				tiles.push_back(Block{{co[0], co[1], co[2]}, 1 << C, vals[(size_t) i], valueMask[(size_t) i] != 0});
		}
	};
	for (uint32_t c = 0; c < nChildren; c++) {
		int o[3];
		origin(o);
		internal(internal, 5, o);
	}
	if ((int64_t) in.pos != desc->blockPos) throw std::runtime_error(path + ": tree topology does not end where the buffers start");
	std::vector<std::vector<float>> leafVals(leaves.size());
	std::vector<std::vector<uint8_t>> leafMask(leaves.size());
	for (size_t l = 0; l < leaves.size(); l++) {
		leafMask[l] = in.mask(512);
		leafVals[l] = vdbValues(in, 512, leafMask[l], maskCompressed, half, bg);
	}
	if ((int64_t) in.pos != desc->endPos) throw std::runtime_error(path + ": leaf buffers do not end at the grid's end");
	// bounding box of the active voxels and active tiles
	for (int a = 0; a < 3; a++) {
		g.lo[a] = INT32_MAX;
		g.hi[a] = INT32_MIN;
	}
	auto grow = [&](const int *p, int span) {
		for (int a = 0; a < 3; a++) {
			g.lo[a] = std::min(g.lo[a], p[a]);
			g.hi[a] = std::max(g.hi[a], p[a] + span - 1);
		}
	};
	for (size_t l = 0; l < leaves.size(); l++)
		for (int i = 0; i < 512; i++)
			if (leafMask[l][(size_t) i]) {
				const int p[3] = {leaves[l].o[0] + (i >> 6), leaves[l].o[1] + ((i >> 3) & 7), leaves[l].o[2] + (i & 7)};
				grow(p, 1);
			}
	for (const Block &t : tiles)
		if (t.active) grow(t.o, t.span);
	if (g.lo[0] > g.hi[0]) throw std::runtime_error(path + ": grid \"" + key + "\" has no active voxels");
	int64_t res[3];
	for (int a = 0; a < 3; a++) {
		res[a] = (int64_t) g.hi[a] - g.lo[a] + 1;
		if (res[a] > 1024) throw std::runtime_error(path + ": active voxels span more than 1024 voxels along an axis");
	}
	g.dense.assign((size_t) (res[0] * res[1] * res[2]), 0.f);
	auto at = [&](int x, int y, int z) -> float & {
		return g.dense[(size_t) ((x - g.lo[0]) + res[0] * ((y - g.lo[1]) + res[1] * (int64_t) (z - g.lo[2])))];
	};
	for (const Block &t : tiles) {
		int a0[3], a1[3]; // the tile clipped to the box
		for (int a = 0; a < 3; a++) {
			a0[a] = std::max(t.o[a], g.lo[a]);
			a1[a] = std::min(t.o[a] + t.span - 1, g.hi[a]);
		}
		for (int z = a0[2]; z <= a1[2]; z++)
			for (int y = a0[1]; y <= a1[1]; y++)
				for (int x = a0[0]; x <= a1[0]; x++) at(x, y, z) = t.v;
	}
	for (size_t l = 0; l < leaves.size(); l++)
		for (int i = 0; i < 512; i++) {
			const int x = leaves[l].o[0] + (i >> 6), y = leaves[l].o[1] + ((i >> 3) & 7), z = leaves[l].o[2] + (i & 7);
			if (x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2]) at(x, y, z) = leafVals[l][(size_t) i];
		}
	for (float v : g.dense)
		if (!(v >= 0.f)) throw std::runtime_error(path + ": negative or NaN densities are not read");
	return g;
}

// The scene graph of KiRaRay's scene JSON (SceneImporter::importNode,
// krrscene.cpp:305-343): "model" is an array, a model file path, or a node
// {"name", "translate", "rotate", "scale", "params"} holding a nested
// "model" or, by "type", a medium (krrscene.cpp:68-143) or a light
// (krrscene.cpp:28-66).  Node transforms compose down the tree.  The
// procedural model types of this build are leaves without a transform.
struct Importer {
	Builder &b;
	krr_scene &s;
	std::string baseDir;
	const std::map<std::string, krr_material> &overrides;
	std::vector<std::pair<int, Json>> bindings; // media bound to meshes by name

	static bool endsWith(const std::string &f, const char *ext) {
		const size_t n = std::strlen(ext);
		if (f.size() < n) return false;
		for (size_t i = 0; i < n; i++)
			if (std::tolower((unsigned char) f[f.size() - n + i]) != ext[i]) return false;
		return true;
	}

	// A model file with the params of the node that names it
	// (SceneImporter::loadModel, krrscene.cpp:8-26).
	void modelFile(const std::string &file, const NodeXform &xf, const Json &params) {
		if (endsWith(file, ".vdb")) {
			vdbVolume(joinPath(baseDir, file), params, xf);
			return;
		}
		if (!endsWith(file, ".obj")) throw std::runtime_error("model file \"" + file + "\": only Wavefront .obj and OpenVDB .vdb files are read");
		loadObj(b, joinPath(baseDir, file), xf, overrides);
	}

	void node(const Json &j, const NodeXform &parent, const Json &parentParams = Json::object()) {
		if (j.isArray()) {
			for (const Json &c : j.items()) node(c, parent, parentParams);
			return;
		}
		if (j.isString()) {
			modelFile(j.asString(), parent, parentParams);
			return;
		}
		if (!j.isObject()) throw std::runtime_error("model entries are arrays, file paths or objects, not " + j.dump());
		const NodeXform xf = parent.then(nodeXform(j));
		const Json params  = j.contains("params") ? j.at("params") : Json::object();
		if (j.contains("model")) {
			node(j.at("model"), xf, params);
			return;
		}
		const std::string t = j.value("type", "mesh");
		if (t == "medium") {
			medium(params, params.value("type", "homogeneous"), xf);
			return;
		}
		if (t == "light") {
			light(params, xf);
			return;
		}
		if (!xf.identity()) throw std::runtime_error("translate / rotate / scale apply to model files, media and nested nodes, not to \"" + t + "\" entries");
		if (t == "cornell_box") cornellBox(b, j);
		else if (t == "icosphere_field") icosphereField(b, j);
		else if (t == "mesh") explicitMesh(b, j);
		else if (t == "quad") quadModel(b, j);
		else throw std::runtime_error("unsupported model type \"" + t + "\"");
	}

	// A medium; "bound": [[lo], [hi]] adds a box without material (a medium
	// interface, material -1) with the medium inside (krrscene.cpp:90-111), "meshes" / "meshes_outside" put
	// it inside / outside the meshes of those names (krrscene.cpp:113-129).
	int medium(const Json &p, const std::string &t, const NodeXform &xf) {
		if (t == "homogeneous") homogeneousMedium(s, p);
		else if (t == "heterogeneous" && p.contains("file")) return vdbVolume(joinPath(baseDir, p.at("file").asString()), p, xf);
		else if (t == "heterogeneous" || t == "value_noise_grid") valueNoiseMedium(s, p);
		else if (t == "dense_grid") denseGridMedium(s, p);
		else throw std::runtime_error("unsupported medium type \"" + t + "\"");
		const int mi = (int) s.media.size() - 1;
		if (p.contains("bound")) {
			const auto &B = p.at("bound").items();
			if (B.size() != 2 || B[0].items().size() != 3 || B[1].items().size() != 3) throw std::runtime_error("medium bound must be [[x, y, z], [x, y, z]]");
			V lo{(float) B[0][0].asNumber(), (float) B[0][1].asNumber(), (float) B[0][2].asNumber()};
			V hi{(float) B[1][0].asNumber(), (float) B[1][1].asNumber(), (float) B[1][2].asNumber()};
			box(lo, hi, xf, mi);
		}
		if (p.contains("meshes") || p.contains("meshes_outside")) bindings.push_back({mi, p});
		return mi;
	}

	// An OpenVDB density grid as a heterogeneous medium inside a box without
	// material around its active voxels (OpenVDBImporter::import,
	// openvdb.cpp:8-67): keys sigma_t (default 1), albedo (0.5), sigma_a /
	// sigma_s (as for homogeneous media), g (0), scale (1, the density
	// multiplier) and key_density ("density").  The box spans the grid's
	// world bounding box, index lo to hi + 1 (NanoVDB's worldBBox), and the
	// node transform places both; this build keeps media axis-aligned, so
	// the transform may translate and scale but not rotate or mirror.
	// Temperature (emission) and albedo grids are not rendered: a file that
	// holds a grid under key_temperature ("temperature") or key_albedo
	// ("albedo") is an error rather than a volume drawn without them.
	int vdbVolume(const std::string &path, const Json &p, const NodeXform &xf) {
		if (xf.L[1] != 0.f || xf.L[2] != 0.f || xf.L[3] != 0.f || xf.L[5] != 0.f || xf.L[6] != 0.f || xf.L[7] != 0.f ||
			!(xf.L[0] > 0.f && xf.L[4] > 0.f && xf.L[8] > 0.f))
			throw std::runtime_error(path + ": OpenVDB volumes may be translated and scaled (by positive factors), not rotated or mirrored");
		const VdbFile f = openVdb(path);
		for (const char *k : {"temperature", "albedo"}) {
			const std::string key = p.value(std::string("key_") + k, k);
			if (vdbHasGrid(f, key)) throw std::runtime_error(path + ": the " + k + " grid \"" + key + "\" is not rendered by this build");
		}
		VdbGrid g = readVdbGrid(f, p.value("key_density", "density"));
		krr_medium m;
		std::memset(&m, 0, sizeof(m));
		m.type		= KRR_MEDIUM_GRID;
		float st[3] = {1, 1, 1}, al[3] = {0.5f, 0.5f, 0.5f};
		p.vec<3>("sigma_t", st);
		p.vec<3>("albedo", al);
		if (p.contains("sigma_a") || p.contains("sigma_s")) {
			float sa[3] = {0.5f, 0.5f, 0.5f}, ss[3] = {0.5f, 0.5f, 0.5f};
			p.vec<3>("sigma_a", sa);
			p.vec<3>("sigma_s", ss);
			for (int a = 0; a < 3; a++) {
				st[a] = sa[a] + ss[a];
				al[a] = ss[a] / st[a];
			}
		}
		V lo, hi;
		float *plo = &lo.x, *phi = &hi.x;
		for (int a = 0; a < 3; a++) {
			m.sigma_t[a] = st[a];
			m.albedo[a]	 = al[a];
			m.res[a]	 = g.hi[a] - g.lo[a] + 1;
			plo[a]		 = (float) (g.origin[a] + g.voxel[a] * g.lo[a]);
			phi[a]		 = (float) (g.origin[a] + g.voxel[a] * ((double) g.hi[a] + 1.0));
		}
		const V wlo = xf.point(lo), whi = xf.point(hi);
		for (int a = 0; a < 3; a++) {
			m.bounds_min[a] = (&wlo.x)[a];
			m.bounds_max[a] = (&whi.x)[a];
		}
		m.g		= (float) p.value("g", 0.0);
		m.scale = (float) p.value("scale", 1.0);
		auto grid = std::make_unique<std::vector<float>>(std::move(g.dense));
		m.density = grid->data();
		s.densities.push_back(std::move(grid));
		s.media.push_back(m);
		const int mi = (int) s.media.size() - 1;
		const size_t slash = path.find_last_of('/');
		box(lo, hi, xf, mi, "Medium box-" + (slash == std::string::npos ? path : path.substr(slash + 1)));
		return mi;
	}

	// Twelve triangles wound so that their normals face out of the box.
	void box(V lo, V hi, const NodeXform &xf, int mi, const std::string &name = "Medium box") {
		const V c{0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
		const V h{0.5f * (hi.x - lo.x), 0.5f * (hi.y - lo.y), 0.5f * (hi.z - lo.z)};
		const V X{h.x, 0, 0}, Y{0, h.y, 0}, Z{0, 0, h.z};
		const V nx{-h.x, 0, 0}, ny{0, -h.y, 0}, nz{0, 0, -h.z};
		// (outward offset, u, v) with u x v along the outward normal
		const V faces[6][3] = {{X, Y, Z}, {nx, Z, Y}, {Y, Z, X}, {ny, X, Z}, {Z, X, Y}, {nz, Y, X}};
		std::vector<V> verts;
		std::vector<int> tris;
		for (const auto &f : faces) {
			const V o = add(c, f[0]);
			const int base = (int) verts.size();
			verts.push_back(xf.point(sub(sub(o, f[1]), f[2])));
			verts.push_back(xf.point(sub(add(o, f[1]), f[2])));
			verts.push_back(xf.point(add(add(o, f[1]), f[2])));
			verts.push_back(xf.point(add(sub(o, f[1]), f[2])));
			for (int k : {0, 1, 2, 0, 2, 3}) tris.push_back(base + k);
		}
		b.addMesh(verts, {}, tris, -1, {0, 0, 0}, mi, -1, name);
	}

	void bindMedia() {
		for (const auto &bd : bindings) {
			for (int side = 0; side < 2; side++) {
				const char *key = side ? "meshes_outside" : "meshes";
				if (!bd.second.contains(key)) continue;
				for (const Json &n : bd.second.at(key).items()) {
					bool found = false;
					for (size_t k = 0; k < s.meshes.size(); k++) {
						if (s.meshNames[k] != n.asString()) continue;
						(side ? s.meshes[k].medium_outside : s.meshes[k].medium_inside) = bd.first;
						found = true;
					}
					if (!found) throw std::runtime_error("medium: no mesh is named \"" + n.asString() + "\"");
				}
			}
		}
	}

	// Lights of the scene graph: infinite lights (constant, or an image
	// "texture" turned by the node's rotation); point, directional and spot
	// lights are not part of this build's light sampling.
	void light(const Json &p, const NodeXform &xf) {
		const std::string t = p.value("type", "infinite");
		if (t != "infinite") throw std::runtime_error("light type \"" + t + "\" is not supported (infinite lights only)");
		float col[3] = {1, 1, 1};
		p.vec<3>("color", col);
		const std::string tex = p.value("texture", "");
		addInfinite(s, col, (float) p.value("scale", 1.0), tex.empty() ? "" : joinPath(baseDir, tex), xf.L);
	}
};

void buildScene(krr_scene &s, const Json &j) {
	Builder b{s};
	// "materials": by index for the procedural models, by name for the
	// materials of model files (krrscene.cpp:198-229)
	std::map<std::string, krr_material> overrides;
	if (j.contains("materials")) {
		for (const Json &m : j.at("materials").items()) {
			krr_material mm = jsonMaterial(m);
			// this build's keys: image files standing for the diffuse colour and the glossiness
			const Json &p = m.contains("params") ? m.at("params") : m;
			for (const char *k : {"diffuse_texture", "specular_texture", "transmission_texture", "emissive_texture", "normal_texture"})
				if (p.contains(k))
					(k[0] == 'd' ? mm.diffuse_texture : k[0] == 's' ? mm.specular_texture : k[0] == 't' ? mm.transmission_texture
					 : k[0] == 'e' ? mm.emissive_texture : mm.normal_texture) = textureFor(s, joinPath(j.value("base_dir", ""), p.value(k, "")));
			// the alpha test's transmission = 1 - opacity (pbrt.cpp:152-156)
			p.vec<3>("transmission", mm.transmission);
			s.materials.push_back(mm);
			if (m.contains("name")) overrides[m.value("name", "")] = mm;
		}
	}
	Importer im{b, s, j.value("base_dir", ""), overrides};
	// "media" before "model": procedural models refer to media by index
	if (j.contains("media")) {
		for (const Json &m : j.at("media").items()) {
			const Json &p = m.contains("params") ? m.at("params") : m;
			const std::string outer = m.value("type", "homogeneous");
			im.medium(p, p.value("type", outer.c_str()), NodeXform{});
		}
	}
	if (j.contains("model")) im.node(j.at("model"), NodeXform{});
	im.bindMedia();
	for (const krr_mesh &m : s.meshes) {
		if (m.material >= (int) s.materials.size()) throw std::runtime_error("mesh material index out of range");
		if (m.medium_inside >= (int) s.media.size() || m.medium_outside >= (int) s.media.size())
			throw std::runtime_error("mesh medium index out of range");
		bool emissive = m.emission[0] > 0 || m.emission[1] > 0 || m.emission[2] > 0;
		if (emissive && !m.has_normals) throw std::runtime_error("emissive meshes need vertex normals (Triangle::sample reads them)");
	}
	if (j.contains("environment")) { // krrscene.cpp:268-274: an image path; or {"tint", "scale", "image", "rotate"}
		const Json &e	 = j.at("environment");
		const std::string base = j.value("base_dir", "");
		float tint[3]	 = {1, 1, 1};
		if (e.isString()) addInfinite(s, tint, 1.f, joinPath(base, e.asString()), nullptr);
		else {
			e.vec<3>("tint", tint);
			const std::string img = e.value("image", "");
			Json rot			  = Json::object();
			if (e.contains("rotate")) rot.set("rotate", e.at("rotate"));
			const NodeXform xf = nodeXform(rot);
			addInfinite(s, tint, (float) e.value("scale", 1.0), img.empty() ? "" : joinPath(base, img), xf.L);
		}
	}
	int w = 0, h = 0;
	if (j.contains("resolution")) {
		const auto &r = j.at("resolution").items();
		if (r.size() != 2) throw std::runtime_error("resolution must be [w, h]");
		w = (int) r[0].asNumber();
		h = (int) r[1].asNumber();
	}
	Json cj		 = j.contains("camera") ? j.at("camera") : Json::object();
	float eye[3] = {0, 1, 3.4f}, target[3] = {0, 1, 0}, up[3] = {0, 1, 0};
	cj.vec<3>("eye", eye);
	cj.vec<3>("target", target);
	cj.vec<3>("up", up);
	if (cj.contains("mData")) { // KiRaRay's camera (and orbit controller)
		const Json ctrl = j.contains("cameraController") ? j.at("cameraController") : Json::object();
		krr_camera c	= referenceCamera(cj, j.contains("cameraController") ? &ctrl : nullptr);
		s.cameras.push_back(c);
		s.viewSizes.push_back({w, h});
	} else if (j.contains("views")) {
		const Json &vj = j.at("views");
		int count	   = (int) vj.value("count", 256.0);
		Rng rng((uint64_t) vj.value("seed", 5.0), 3);
		float radius	= (float) vj.value("radius", 2.0);
		float center[3] = {0.5f, 0.5f, 0.5f};
		vj.vec<3>("center", center);
		std::vector<int> sizes = {128, 256, 512, 1024};
		if (vj.contains("sizes")) {
			sizes.clear();
			for (auto &x : vj.at("sizes").items()) sizes.push_back((int) x.asNumber());
		}
		for (int v = 0; v < count; v++) {
			// uniform direction on the sphere, upper hemisphere biased away from the ground
			float z	  = 1.f - 2.f * rng.uniform();
			float phi = 6.2831853f * rng.uniform();
			float rr  = std::sqrt(std::max(0.f, 1.f - z * z));
			V dir	  = {rr * std::cos(phi), std::fabs(z) * 0.8f + 0.1f, rr * std::sin(phi)};
			dir		  = nrmz(dir);
			int side  = sizes[rng.next() % sizes.size()];
			V c		  = {center[0], center[1], center[2]};
			s.cameras.push_back(lookAtCamera(add(c, mul(dir, radius)), c, {0, 1, 0}, side, side, cj));
			s.viewSizes.push_back({side, side});
		}
	} else {
		s.cameras.push_back(lookAtCamera({eye[0], eye[1], eye[2]}, {target[0], target[1], target[2]}, {up[0], up[1], up[2]}, w, h, cj));
		s.viewSizes.push_back({w, h});
	}
	for (const krr_camera &c : s.cameras)
		if (c.medium >= (int) s.media.size()) throw std::runtime_error("camera medium index out of range");
}

} // namespace
} // namespace krr

extern "C" {

int krr_scene_create(const char *json, krr_scene **out) {
	krr::clearError();
	if (!json || !out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_scene_create: null argument");
	*out = nullptr;
	try {
		krr::Json j = krr::Json::parse(json);
		if (!j.isObject()) return krr::fail(KRR_ERR_BAD_CONFIG, "scene config must be a JSON object");
		auto s		= std::make_unique<krr_scene>();
		krr::buildScene(*s, j);
		*out = s.release();
		return KRR_OK;
	} catch (const std::bad_alloc &) {
		return krr::fail(KRR_ERR_OUT_OF_MEMORY, "krr_scene_create: out of host memory");
	} catch (const std::exception &e) {
		return krr::fail(KRR_ERR_BAD_CONFIG, "krr_scene_create: %s", e.what());
	}
}

int krr_scene_get_desc(const krr_scene *s, krr_scene_desc *d) {
	if (!s || !d) return krr::fail(KRR_ERR_INVALID_ARG, "krr_scene_get_desc: null argument");
	std::memset(d, 0, sizeof(*d));
	d->n_vertices		  = (int64_t) (s->positions.size() / 3);
	d->positions		  = s->positions.data();
	d->normals			  = s->anyNormals ? s->normals.data() : nullptr;
	d->n_triangles		  = (int64_t) (s->indices.size() / 3);
	d->indices			  = s->indices.data();
	d->n_meshes			  = (int32_t) s->meshes.size();
	d->meshes			  = s->meshes.data();
	d->n_materials		  = (int32_t) s->materials.size();
	d->materials		  = s->materials.data();
	d->n_media			  = (int32_t) s->media.size();
	d->media			  = s->media.data();
	d->n_infinite_lights  = (int32_t) s->infinite.size();
	d->infinite_lights	  = s->infinite.data();
	d->texcoords		  = s->anyTexcoords ? s->texcoords.data() : nullptr;
	d->n_textures		  = (int32_t) s->textures.size();
	d->textures			  = s->textures.data();
	d->tangents			  = s->anyTangents ? s->tangents.data() : nullptr;
	return KRR_OK;
}

int krr_scene_get_camera(const krr_scene *s, int view, krr_camera *out) {
	if (!s || !out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_scene_get_camera: null argument");
	if (view < 0 || view >= (int) s->cameras.size()) return krr::fail(KRR_ERR_INVALID_ARG, "krr_scene_get_camera: view %d out of range", view);
	*out = s->cameras[view];
	return KRR_OK;
}

int krr_scene_view_count(const krr_scene *s) { return s ? (int) s->cameras.size() : 0; }

int krr_scene_view_size(const krr_scene *s, int view, int32_t *w, int32_t *h) {
	if (!s || !w || !h) return krr::fail(KRR_ERR_INVALID_ARG, "krr_scene_view_size: null argument");
	if (view < 0 || view >= (int) s->viewSizes.size()) return krr::fail(KRR_ERR_INVALID_ARG, "krr_scene_view_size: view out of range");
	*w = s->viewSizes[view].first;
	*h = s->viewSizes[view].second;
	return KRR_OK;
}

int krr_scene_destroy(krr_scene *s) {
	delete s;
	return KRR_OK;
}

} // extern "C"
