// wavefront_pass.cpp - WavefrontPathTracer hooks forwarded to the C ABI.
#include "wavefront_pass.h"

#include <vector>

namespace krr {

KRR_REGISTER_PASS_DEF(WavefrontPathTracer);

namespace {
const char *const kReferenceKeys[] = {"nee", "enable_medium", "max_depth", "rr", "enable_clamp", "clamp_max"};

bool isReferenceKey(const std::string &k) {
	for (const char *r : kReferenceKeys)
		if (k == r) return true;
	return false;
}
} // namespace

// integrator.h:85-94
void to_json(json &j, const WavefrontPathTracer &p) {
	j = json::object();
	for (const auto &kv : p.extraParams.members()) j.set(kv.first, kv.second);
	j.set("nee", json(p.enableNEE));
	j.set("enable_medium", json(p.enableMedium));
	j.set("max_depth", json((double) p.maxDepth));
	j.set("rr", json((double) p.probRR));
	j.set("enable_clamp", json(p.enableClamp));
	j.set("clamp_max", json((double) p.clampMax));
}

// integrator.h:96-103 (same keys and defaults); other keys are kept for the
// library, which validates them when the pass is created.
void from_json(const json &j, WavefrontPathTracer &p) {
	try {
		if (!j.isObject() && !j.isNull()) throw std::runtime_error("params must be a JSON object");
		p.enableNEE	   = j.value("nee", true);
		p.enableMedium = j.value("enable_medium", true);
		p.maxDepth	   = (int) j.value("max_depth", 10.0);
		p.probRR	   = (float) j.value("rr", 0.8);
		p.enableClamp  = j.value("enable_clamp", false);
		p.clampMax	   = (float) j.value("clamp_max", 1e3);
		p.extraParams  = json::object();
		if (j.isObject())
			for (const auto &kv : j.members())
				if (!isReferenceKey(kv.first)) p.extraParams.set(kv.first, kv.second);
	} catch (const Error &) {
		throw;
	} catch (const std::exception &e) {
		throw Error(KRR_ERR_BAD_CONFIG, std::string("WavefrontPathTracer params: ") + e.what());
	}
}

WavefrontPathTracer::~WavefrontPathTracer() {
	if (mWf) krr_wf_destroy(mWf);
}

void WavefrontPathTracer::ensureCreated() {
	if (mWf) return;
	json j;
	to_json(j, *this);
	KRR_CALL(krr_wf_create(j.dump().c_str(), &mWf));
}

// integrator.cpp:22-48: the queues are allocated by resize; creating the
// library pass validates the params.
void WavefrontPathTracer::initialize() {
	ensureCreated();
	uploadSceneIfNeeded();
	if (mSized) KRR_CALL(krr_wf_resize(mWf, mSize[0], mSize[1]));
}

void WavefrontPathTracer::finalize() {
	if (mWf) {
		KRR_CALL(krr_wf_destroy(mWf));
		mWf = nullptr;
	}
}

// integrator.cpp:181-184
void WavefrontPathTracer::resize(const Vector2i &size) {
	mSize  = size;
	mSized = true;
	if (mWf) KRR_CALL(krr_wf_resize(mWf, size[0], size[1]));
}

// integrator.cpp:186-203
void WavefrontPathTracer::setScene(Scene::SharedPtr scene) {
	mScene		= scene;
	mSceneDirty = true;
	if (mWf) uploadSceneIfNeeded();
}

void WavefrontPathTracer::uploadSceneIfNeeded() {
	if (!mSceneDirty || !mScene) return;
	KRR_CALL(krr_wf_set_scene(mWf, &mScene->getDesc()));
	mSceneDirty = false;
}

// integrator.cpp:205-221
void WavefrontPathTracer::beginFrame(RenderContext *context) {
	if (!mEnable) return;
	ensureCreated();
	uploadSceneIfNeeded();
	if (!mScene) throw Error(KRR_ERR_NO_SCENE, "WavefrontPathTracer::beginFrame: no scene set");
	const krr_camera cam = mScene->getCamera(0);
	KRR_CALL(krr_wf_begin_frame(mWf, &cam, (uint64_t) getFrameIndex()));
}

// integrator.cpp:223-267: writes the context's colour texture.
void WavefrontPathTracer::render(RenderContext *context) {
	if (!mEnable) return;
	ensureCreated();
	const CudaRenderTarget t = context->getColorTexture()->getCudaRenderTarget();
	if (t.width != mSize[0] || t.height != mSize[1])
		throw Error(KRR_ERR_INVALID_ARG, "WavefrontPathTracer::render: colour texture size differs from the pass size");
	KRR_CALL(krr_wf_render(mWf, t.data, context->getCudaStream()));
}

void WavefrontPathTracer::endFrame(RenderContext *context) {
	if (mWf) KRR_CALL(krr_wf_end_frame(mWf));
}

krr_stats WavefrontPathTracer::getStats() {
	krr_stats s{};
	if (mWf) KRR_CALL(krr_wf_stats(mWf, &s));
	return s;
}

json WavefrontPathTracer::getParams() const {
	if (!mWf) {
		json j;
		to_json(j, *this);
		return j;
	}
	const int n = krr_wf_get_params(mWf, nullptr, 0);
	checkStatus(n, "krr_wf_get_params");
	std::vector<char> buf((size_t) n + 1);
	KRR_CALL(krr_wf_get_params(mWf, buf.data(), (int32_t) buf.size()));
	return json::parse(buf.data());
}

} // namespace krr
