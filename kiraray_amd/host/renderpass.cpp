// renderpass.cpp - factory, scene, render context and pass helpers of the
// host-side plugin interface (renderpass.h).
#include "renderpass.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>

namespace krr {

void checkStatus(int status, const char *call) {
	if (status >= 0) return;
	const char *msg = krr_last_error();
	throw Error(status, std::string(call) + " failed (" + std::to_string(status) + "): " + (msg ? msg : ""));
}

static void checkHip(hipError_t e, const char *call) {
	if (e != hipSuccess) throw Error(KRR_ERR_HIP, std::string(call) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- Scene

Scene::SharedPtr Scene::fromJson(const std::string &sceneJson) {
	std::shared_ptr<Scene> s(new Scene());
	KRR_CALL(krr_scene_create(sceneJson.c_str(), &s->mHandle));
	KRR_CALL(krr_scene_get_desc(s->mHandle, &s->mDesc));
	const int n = krr_scene_view_count(s->mHandle);
	for (int v = 0; v < n; v++) {
		krr_camera c{};
		KRR_CALL(krr_scene_get_camera(s->mHandle, v, &c));
		s->mCameras.push_back(c);
	}
	return s;
}

Scene::SharedPtr Scene::fromDesc(const krr_scene_desc &desc, std::vector<krr_camera> cameras) {
	if (cameras.empty()) throw Error(KRR_ERR_INVALID_ARG, "Scene::fromDesc: at least one camera is needed");
	std::shared_ptr<Scene> s(new Scene());
	s->mDesc	= desc;
	s->mCameras = std::move(cameras);
	return s;
}

Scene::~Scene() {
	if (mHandle) krr_scene_destroy(mHandle);
}

krr_camera Scene::getCamera(int view) const {
	if (view < 0 || view >= (int) mCameras.size()) throw Error(KRR_ERR_INVALID_ARG, "Scene::getCamera: view out of range");
	return mCameras[view];
}

int Scene::getViewCount() const { return (int) mCameras.size(); }

void Scene::setAspectRatio(float aspectRatio) {
	for (auto &c : mCameras) c.aspect_ratio = aspectRatio;
}

Vector2i Scene::getResolution(int view) const {
	int32_t w = 0, h = 0;
	if (mHandle) KRR_CALL(krr_scene_view_size(mHandle, view, &w, &h));
	return {w, h};
}

// -------------------------------------------------------- RenderContext

RenderContext::~RenderContext() {
	if (mColor.mTarget.data) (void) hipFree(mColor.mTarget.data);
	if (mStream) (void) hipStreamDestroy((hipStream_t) mStream);
}

void *RenderContext::getCudaStream() {
	if (!mStream) {
		hipStream_t st = nullptr;
		checkHip(hipStreamCreate(&st), "hipStreamCreate"); // default priority
		mStream = st;
	}
	return mStream;
}

void RenderContext::resize(const Vector2i size) {
	if (size[0] < 0 || size[1] < 0) throw Error(KRR_ERR_INVALID_ARG, "RenderContext::resize: negative size");
	mSize = size;
}

// (Re)allocates the colour texture when the context size changed.
RenderTexture *RenderContext::getColorTexture() {
	if (mColor.getSize() == mSize) return &mColor;
	hipStream_t st = (hipStream_t) getCudaStream();
	checkHip(hipStreamSynchronize(st), "hipStreamSynchronize");
	if (mColor.mTarget.data) checkHip(hipFree(mColor.mTarget.data), "hipFree");
	mColor.mTarget	   = CudaRenderTarget{};
	const size_t bytes = (size_t) mSize[0] * mSize[1] * 4 * sizeof(float);
	void *p			   = nullptr;
	if (bytes) {
		checkHip(hipMalloc(&p, bytes), "hipMalloc");
		checkHip(hipMemsetAsync(p, 0, bytes, st), "hipMemsetAsync");
	}
	mColor.mTarget = CudaRenderTarget{(float *) p, mSize[0], mSize[1]};
	return &mColor;
}

std::vector<float> RenderContext::readColor() {
	const CudaRenderTarget t = getColorTexture()->getCudaRenderTarget();
	std::vector<float> out((size_t) t.width * t.height * 4);
	checkHip(hipStreamSynchronize((hipStream_t) getCudaStream()), "hipStreamSynchronize");
	if (!out.empty()) checkHip(hipMemcpy(out.data(), t.data, out.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
	return out;
}

// ------------------------------------------------------------ RenderPass

size_t RenderPass::getFrameIndex() const {
	if (!mDeviceManager) throw Error(KRR_ERR_INVALID_ARG, getName() + ": no host loop attached (setDeviceManager)");
	return mDeviceManager->getFrameIndex();
}

Vector2i RenderPass::getFrameSize() const {
	if (!mDeviceManager) throw Error(KRR_ERR_INVALID_ARG, getName() + ": no host loop attached (setDeviceManager)");
	return mDeviceManager->getFrameSize();
}

// ------------------------------------------------------------- Factory

RenderPassFactory::map_type &RenderPassFactory::getMap() {
	static map_type m;
	return m;
}

RenderPassFactory::configured_map_type &RenderPassFactory::getConfiguredMap() {
	static configured_map_type m;
	return m;
}

RenderPass::SharedPtr RenderPassFactory::createInstance(const std::string &name) {
	auto it = getMap().find(name);
	if (it == getMap().end()) {
		std::fprintf(stderr, "[krr] could not create instance for %s: check if the pass is registered.\n", name.c_str());
		return nullptr;
	}
	return it->second();
}

RenderPass::SharedPtr RenderPassFactory::deserizeInstance(const std::string &name, const json &serde) {
	auto it = getConfiguredMap().find(name);
	if (it == getConfiguredMap().end()) {
		std::fprintf(stderr, "[krr] could not deserialize instance for %s: check if the pass is registered.\n", name.c_str());
		return nullptr;
	}
	return it->second(serde);
}

std::vector<std::string> RenderPassFactory::registeredNames() {
	std::vector<std::string> names;
	for (const auto &kv : getMap()) names.push_back(kv.first);
	return names;
}

} // namespace krr
