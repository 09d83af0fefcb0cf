// megakernel_pass.cpp - MegakernelPathTracer hooks forwarded to the C ABI.
#include "megakernel_pass.h"

#include <vector>

namespace krr {

KRR_REGISTER_PASS_DEF(MegakernelPathTracer);

// pathtracer.h:37-43
void to_json(json &j, const MegakernelPathTracer &p) {
	j = json::object();
	for (const auto &kv : p.extraParams.members()) j.set(kv.first, kv.second);
	j.set("nee", json(p.enableNEE));
	j.set("max_depth", json((double) p.maxDepth));
	j.set("rr", json((double) p.probRR));
}

// pathtracer.h:45-49 (same keys and defaults); other keys are kept for the
// library, which validates them when the pass is created.
void from_json(const json &j, MegakernelPathTracer &p) {
	try {
		if (!j.isObject() && !j.isNull()) throw std::runtime_error("params must be a JSON object");
		p.enableNEE	  = j.value("nee", true);
		p.maxDepth	  = (int) j.value("max_depth", 10.0);
		p.probRR	  = (float) j.value("rr", 0.8);
		p.extraParams = json::object();
		if (j.isObject())
			for (const auto &kv : j.members())
				if (kv.first != "nee" && kv.first != "max_depth" && kv.first != "rr") p.extraParams.set(kv.first, kv.second);
	} catch (const Error &) {
		throw;
	} catch (const std::exception &e) {
		throw Error(KRR_ERR_BAD_CONFIG, std::string("MegakernelPathTracer params: ") + e.what());
	}
}

MegakernelPathTracer::~MegakernelPathTracer() {
	if (mMk) krr_mk_destroy(mMk);
}

void MegakernelPathTracer::ensureCreated() {
	if (mMk) return;
	json j;
	to_json(j, *this);
	KRR_CALL(krr_mk_create(j.dump().c_str(), &mMk));
}

void MegakernelPathTracer::initialize() {
	ensureCreated();
	uploadSceneIfNeeded();
	if (mSized) KRR_CALL(krr_mk_resize(mMk, mSize[0], mSize[1]));
}

void MegakernelPathTracer::finalize() {
	if (mMk) {
		KRR_CALL(krr_mk_destroy(mMk));
		mMk = nullptr;
	}
}

// pathtracer.cpp:20-23
void MegakernelPathTracer::resize(const Vector2i &size) {
	mSize  = size;
	mSized = true;
	if (mMk) KRR_CALL(krr_mk_resize(mMk, size[0], size[1]));
}

// pathtracer.cpp:25-39
void MegakernelPathTracer::setScene(Scene::SharedPtr scene) {
	mScene		= scene;
	mSceneDirty = true;
	if (mMk) uploadSceneIfNeeded();
}

void MegakernelPathTracer::uploadSceneIfNeeded() {
	if (!mSceneDirty || !mScene) return;
	KRR_CALL(krr_mk_set_scene(mMk, &mScene->getDesc()));
	mSceneDirty = false;
}

// pathtracer.cpp:52-69: the camera and frame index of this frame, then one
// launch that writes the context's colour texture.
void MegakernelPathTracer::render(RenderContext *context) {
	if (!mEnable) return;
	ensureCreated();
	uploadSceneIfNeeded();
	if (!mScene) throw Error(KRR_ERR_NO_SCENE, "MegakernelPathTracer::render: no scene set");
	const CudaRenderTarget t = context->getColorTexture()->getCudaRenderTarget();
	if (t.width != mSize[0] || t.height != mSize[1])
		throw Error(KRR_ERR_INVALID_ARG, "MegakernelPathTracer::render: colour texture size differs from the pass size");
	const krr_camera cam = mScene->getCamera(0);
	KRR_CALL(krr_mk_begin_frame(mMk, &cam, (uint64_t) getFrameIndex()));
	KRR_CALL(krr_mk_render(mMk, t.data, context->getCudaStream()));
}

krr_stats MegakernelPathTracer::getStats() {
	krr_stats s{};
	if (mMk) KRR_CALL(krr_mk_stats(mMk, &s));
	return s;
}

json MegakernelPathTracer::getParams() const {
	if (!mMk) {
		json j;
		to_json(j, *this);
		return j;
	}
	const int n = krr_mk_get_params(mMk, nullptr, 0);
	checkStatus(n, "krr_mk_get_params");
	std::vector<char> buf((size_t) n + 1);
	KRR_CALL(krr_mk_get_params(mMk, buf.data(), (int32_t) buf.size()));
	return json::parse(buf.data());
}

} // namespace krr
