// host_tests.cpp - checks of the C++ render-pass layer (kiraray_amd/host/),
// driven by tests/test_host_cpp.py.
//
//   host_tests factory          registration, createInstance, deserialize
//   host_tests params           to_json / from_json round trip and errors
//   host_tests config           RenderApp::loadConfig without a device
//   host_tests render W H F OUT Cornell box through RenderApp for F frames;
//                               writes the colour texture (float32 RGBA,
//                               row-major as stored) to OUT and prints the
//                               ray counts of the last frame
//   host_tests render_mk W H F OUT  the same through MegakernelPathTracer
//   host_tests measure F REF OUT IMG  Cornell 256x256 through the passes
//                               [WavefrontPathTracer, AccumulatePass,
//                               ErrorMeasurePass(REF .npy, every 4 frames,
//                               saved to OUT)] for F frames; writes the
//                               accumulated colour texture to IMG
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "megakernel_pass.h"
#include "passes.h"
#include "render_app.h"
#include "wavefront_pass.h"

using namespace krr;

static int gFailures = 0;
#define EXPECT(cond)                                                                                                   \
	do {                                                                                                               \
		if (!(cond)) {                                                                                                 \
			std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #cond);                     \
			gFailures++;                                                                                               \
		}                                                                                                              \
	} while (0)

static const char *kCornell = R"({"model": [{"type": "cornell_box"}],
	"camera": {"eye": [0, 1, 2.8], "target": [0, 1, 0]}})";

static int testFactory() {
	bool found = false;
	for (const auto &n : RenderPassFactory::registeredNames()) found |= n == "WavefrontPathTracer";
	EXPECT(found);
	auto p = RenderPassFactory::createInstance("WavefrontPathTracer");
	EXPECT(p != nullptr);
	EXPECT(p && p->getName() == "WavefrontPathTracer");
	EXPECT(p && p->enabled() && p->isCudaPass());
	EXPECT(RenderPassFactory::createInstance("AccumulatePass") != nullptr);
	EXPECT(RenderPassFactory::createInstance("ErrorMeasurePass") != nullptr);
	EXPECT(RenderPassFactory::createInstance("ToneMappingPass") == nullptr);
	EXPECT(RenderPassFactory::deserizeInstance("NoSuchPass", json::object()) == nullptr);
	auto q = RenderPassFactory::deserizeInstance("WavefrontPathTracer", json::parse(R"({"max_depth": 5, "nee": false})"));
	auto *w = dynamic_cast<WavefrontPathTracer *>(q.get());
	EXPECT(w != nullptr);
	EXPECT(w && w->maxDepth == 5 && !w->enableNEE && w->enableMedium && std::fabs(w->probRR - 0.8f) < 1e-7f);
	// MegakernelPathTracer: the reference's keys and defaults (pathtracer.h:45-49)
	auto m	 = RenderPassFactory::deserizeInstance("MegakernelPathTracer", json::parse(R"({"rr": 0.5, "spp": 2})"));
	auto *mk = dynamic_cast<MegakernelPathTracer *>(m.get());
	EXPECT(mk && mk->getName() == "MegakernelPathTracer");
	EXPECT(mk && mk->enableNEE && mk->maxDepth == 10 && std::fabs(mk->probRR - 0.5f) < 1e-7f);
	EXPECT(mk && mk->getParams().value("spp", 0.0) == 2.0);
	EXPECT(mk && mk->getParams().value("rr", 0.0) == 0.5);
	return 0;
}

static int testParams() {
	WavefrontPathTracer p;
	// integrator.h:96-103 defaults
	EXPECT(p.enableNEE && p.enableMedium && p.maxDepth == 10 && !p.enableClamp);
	EXPECT(std::fabs(p.probRR - 0.8f) < 1e-7f && p.clampMax == 1e3f);
	from_json(json::parse(R"({"rr": 0.5, "spp": 4, "rr_min_depth": 3, "clamp_max": 7})"), p);
	json j;
	to_json(j, p);
	EXPECT(std::fabs(j.value("rr", 0.0) - 0.5) < 1e-7);
	EXPECT(j.value("spp", 0.0) == 4.0 && j.value("rr_min_depth", 0.0) == 3.0);
	EXPECT(j.value("clamp_max", 0.0) == 7.0 && j.value("max_depth", 0.0) == 10.0);
	WavefrontPathTracer r;
	from_json(j, r);
	json j2;
	to_json(j2, r);
	EXPECT(j.dump() == j2.dump());
	// a wrongly typed key is a configuration error
	int code = 0;
	try {
		from_json(json::parse(R"({"max_depth": "ten"})"), r);
	} catch (const Error &e) {
		code = e.code();
	}
	EXPECT(code == KRR_ERR_BAD_CONFIG);
	// AccumulatePass / ErrorMeasurePass: the reference's enum strings
	AccumulatePass a;
	from_json(json::parse(R"({"spp": 64, "mode": "moving average", "precision": "double"})"), a);
	json ja;
	to_json(ja, a);
	EXPECT(ja.value("spp", 0.0) == 64.0 && ja.value("mode", "") == "moving average" && ja.value("precision", "") == "double");
	from_json(json::parse(R"({"mode": 0, "precision": 1})"), a);
	to_json(ja, a);
	EXPECT(ja.value("mode", "") == "accumulate" && ja.value("precision", "") == "double" && ja.value("spp", 1.0) == 0.0);
	ErrorMeasurePass e;
	from_json(json::parse(R"({"metric": "mse", "continuous": true, "interval": 8})"), e);
	json je;
	to_json(je, e);
	EXPECT(je.value("metric", "") == "mse" && je.value("interval", 0.0) == 8.0 && je.value("continuous", false));
	code = 0;
	try {
		from_json(json::parse(R"({"interval": 0})"), e);
	} catch (const Error &err) {
		code = err.code();
	}
	EXPECT(code == KRR_ERR_BAD_CONFIG);
	code = 0;
	try {
		from_json(json::parse(R"({"mode": 5})"), a);
	} catch (const Error &err) {
		code = err.code();
	}
	EXPECT(code == KRR_ERR_BAD_CONFIG);
	return 0;
}

static int testConfig() {
	RenderApp app;
	app.loadConfig(json::parse(std::string(R"({"resolution": [96, 64], "passes": [
		{"enable": true, "name": "WavefrontPathTracer", "params": {"nee": true, "rr": 0.8, "max_depth": 6}},
		{"enable": true, "name": "AccumulatePass", "params": {"spp": 0, "mode": "accumulate"}},
		{"enable": false, "name": "DenoisePass"}],
		"scene": )") + kCornell + "}"));
	EXPECT(app.getRenderPasses().size() == 2);
	EXPECT(app.getRenderPasses().size() == 2 && app.getRenderPasses()[1]->getName() == "AccumulatePass");
	EXPECT(app.getFrameSize() == Vector2i(96, 64));
	EXPECT(app.getScene() && app.getScene()->getDesc().n_triangles == 36);
	auto *w = dynamic_cast<WavefrontPathTracer *>(app.getRenderPasses()[0].get());
	EXPECT(w && w->maxDepth == 6);
	int code = 0;
	try {
		app.loadConfig(json::parse(R"({"resolution": [1, 2, 3]})"));
	} catch (const Error &e) {
		code = e.code();
	}
	EXPECT(code == KRR_ERR_BAD_CONFIG);
	// the pass list of the reference's example_megakernel.json
	RenderApp mapp;
	mapp.loadConfig(json::parse(std::string(R"({"resolution": [750, 750], "passes": [
		{"enable": true, "name": "MegakernelPathTracer", "params": {"nee": true, "rr": 0.8, "max_depth": 10}},
		{"enable": true, "name": "AccumulatePass", "params": {"spp": 0, "mode": "accumulate"}},
		{"enable": false, "name": "DenoisePass"},
		{"enable": true, "name": "ToneMappingPass", "params": {"exposure": 0.5, "operator": "aces"}}],
		"scene": )") + kCornell + "}"));
	EXPECT(mapp.getRenderPasses().size() == 2 && mapp.getRenderPasses()[0]->getName() == "MegakernelPathTracer");
	return 0;
}

template <class Pass> static int testRender(int w, int h, int frames, const char *out) {
	RenderApp app;
	auto pass = std::make_shared<Pass>();
	from_json(json::parse(R"({"max_depth": 8, "spp": 2})"), *pass);
	app.addRenderPassToBack(pass);
	app.setScene(Scene::fromJson(kCornell));
	app.resize({w, h});
	app.initialize();
	app.run(frames);
	const std::vector<float> img = app.getRenderContext()->readColor();
	const krr_stats st			  = pass->getStats();
	FILE *f						  = std::fopen(out, "wb");
	if (!f) return 2;
	std::fwrite(img.data(), sizeof(float), img.size(), f);
	std::fclose(f);
	std::printf("{\"frame_index\": %zu, \"closest_rays\": %llu, \"shadow_rays\": %llu}\n", app.getFrameIndex(),
				(unsigned long long) st.closest_rays, (unsigned long long) st.shadow_rays);
	app.finalize();
	return 0;
}

static int testMeasure(int frames, const char *ref, const char *out, const char *imgOut) {
	RenderApp app;
	app.loadConfig(json::parse(std::string(R"({"resolution": [256, 256], "passes": [
		{"enable": true, "name": "WavefrontPathTracer", "params": {"max_depth": 8, "spp": 1}},
		{"enable": true, "name": "AccumulatePass", "params": {"spp": 0, "mode": "accumulate", "precision": "double"}},
		{"enable": true, "name": "ErrorMeasurePass", "params": {"metric": "rel_mse", "continuous": true, "interval": 4,
		 "save": true, "reference": ")") + ref + R"(", "output": ")" + out + R"("}}],
		"scene": )" + kCornell + "}"));
	app.initialize();
	app.run(frames);
	const std::vector<float> img = app.getRenderContext()->readColor();
	FILE *f						  = std::fopen(imgOut, "wb");
	if (!f) return 2;
	std::fwrite(img.data(), sizeof(float), img.size(), f);
	std::fclose(f);
	auto *acc = dynamic_cast<AccumulatePass *>(app.getRenderPasses()[1].get());
	EXPECT(acc && acc->getAccumCount() == (size_t) frames);
	app.finalize(); // writes OUT
	return 0;
}

int main(int argc, char **argv) {
	if (argc < 2) {
		std::fprintf(stderr, "usage: host_tests factory|params|config|render W H FRAMES OUT\n");
		return 2;
	}
	const std::string cmd = argv[1];
	int rc				  = 0;
	try {
		if (cmd == "factory") rc = testFactory();
		else if (cmd == "params") rc = testParams();
		else if (cmd == "config") rc = testConfig();
		else if (cmd == "render" && argc == 6)
			rc = testRender<WavefrontPathTracer>(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
		else if (cmd == "render_mk" && argc == 6)
			rc = testRender<MegakernelPathTracer>(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
		else if (cmd == "measure" && argc == 6) rc = testMeasure(std::atoi(argv[2]), argv[3], argv[4], argv[5]);
		else {
			std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
			return 2;
		}
	} catch (const Error &e) {
		std::fprintf(stderr, "krr::Error %d: %s\n", e.code(), e.what());
		return 3;
	}
	if (gFailures) std::fprintf(stderr, "%d expectation(s) failed\n", gFailures);
	return rc ? rc : (gFailures ? 1 : 0);
}
