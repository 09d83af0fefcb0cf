// renderpass.h - host-side plugin interface of the gfx950 wavefront tracer.
//
// The same hooks, names and call order as KiRaRay's RenderPass
// (src/core/renderpass.h:138-200), its factory and registration macros
// (src/core/renderpass.h:202-273) and the headless part of RenderContext
// (src/core/renderpass.h:96-136).  The render target is a linear W x H
// float4 (RGBA32F) buffer in device memory; the stream is a HIP stream at
// default priority.
//
// Errors: the reference logs Fatal and exits (src/util/check.h:8-42).  Here
// every failed library call throws krr::Error; left uncaught it still ends
// the process, and callers that want to recover can catch it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.h"
#include "krr_wavefront.h"

namespace krr {

using json = Json;

class Error : public std::runtime_error {
public:
	Error(int code, const std::string &what) : std::runtime_error(what), mCode(code) {}
	int code() const { return mCode; }

private:
	int mCode;
};

// Throws krr::Error for a negative status of the C ABI, with krr_last_error().
void checkStatus(int status, const char *call);
#define KRR_CALL(expr) ::krr::checkStatus((expr), #expr)

struct Vector2i {
	int v[2] = {0, 0};
	Vector2i() = default;
	Vector2i(int x, int y) : v{x, y} {}
	int operator[](int i) const { return v[i]; }
	int &operator[](int i) { return v[i]; }
	int x() const { return v[0]; }
	int y() const { return v[1]; }
	bool operator==(const Vector2i &o) const { return v[0] == o.v[0] && v[1] == o.v[1]; }
	bool operator!=(const Vector2i &o) const { return !(*this == o); }
};

// Scene: the scene description the pass uploads (Scene::SharedPtr of the
// reference).  Built from the procedural scene JSON of krr_scene_create, or
// wrapped around a caller-owned krr_scene_desc + cameras.
class Scene {
public:
	using SharedPtr = std::shared_ptr<Scene>;
	static SharedPtr fromJson(const std::string &sceneJson);
	static SharedPtr fromDesc(const krr_scene_desc &desc, std::vector<krr_camera> cameras);
	~Scene();
	const krr_scene_desc &getDesc() const { return mDesc; }
	krr_camera getCamera(int view = 0) const;
	int getViewCount() const;
	// Camera::setAspectRatio (src/core/camera.h:115) on every view.
	void setAspectRatio(float aspectRatio);
	// Resolution stored with the scene ("resolution" key), (0, 0) if none.
	Vector2i getResolution(int view = 0) const;

private:
	Scene() = default;
	krr_scene *mHandle = nullptr;
	krr_scene_desc mDesc{};
	std::vector<krr_camera> mCameras;
};

// Device render target (CudaRenderTarget counterpart, src/core/device/cuda.h:33-45):
// pixel (x, y) of the image is row H-1-y of `data`.
struct CudaRenderTarget {
	float *data = nullptr; // W * H * 4 floats
	int width = 0, height = 0;
};

class RenderTexture {
public:
	CudaRenderTarget getCudaRenderTarget() const { return mTarget; }
	Vector2i getSize() const { return {mTarget.width, mTarget.height}; }

private:
	friend class RenderContext;
	CudaRenderTarget mTarget;
};

// Headless RenderContext: owns the colour texture (device float4 buffer)
// and the HIP stream every pass launches on.  Both are created on first
// use, so a context can be configured without a device.
class RenderContext {
public:
	using SharedPtr = std::shared_ptr<RenderContext>;
	RenderContext() = default;
	~RenderContext();
	RenderContext(const RenderContext &)			= delete;
	RenderContext &operator=(const RenderContext &) = delete;
	RenderTexture *getColorTexture();
	void *getCudaStream(); // hipStream_t
	Scene::SharedPtr getScene() const { return mScene; }
	void setScene(Scene::SharedPtr scene) { mScene = scene; }
	void resize(const Vector2i size);
	Vector2i getSize() const { return mSize; }
	// Copies the colour texture to the host (waits for the stream).
	std::vector<float> readColor();

private:
	RenderTexture mColor;
	Vector2i mSize{0, 0};
	void *mStream = nullptr;
	Scene::SharedPtr mScene;
};

// Source of the host loop's frame index and back-buffer size
// (DeviceManager, src/core/window.cpp:457,487-490).
class DeviceManager {
public:
	virtual ~DeviceManager()				  = default;
	virtual size_t getFrameIndex() const	  = 0;
	virtual Vector2i getFrameSize() const	  = 0;
};

class RenderPass {
public:
	using SharedPtr = std::shared_ptr<RenderPass>;
	RenderPass()	= default;
	virtual ~RenderPass() = default;

	virtual void resizing() {}
	virtual void resize(const Vector2i &size) {}
	virtual void setEnable(bool enable) { mEnable = enable; }
	virtual void setScene(Scene::SharedPtr scene) { mScene = scene; }
	virtual Scene::SharedPtr getScene() { return mScene; }
	virtual void setDeviceManager(DeviceManager *deviceManager) { mDeviceManager = deviceManager; }

	virtual void tick(float elapsedSeconds) {}
	virtual void beginFrame(RenderContext *context) {}
	virtual void render(RenderContext *context) {}
	virtual void endFrame(RenderContext *context) {}
	virtual void renderUI() {}

	virtual void initialize() {}
	virtual void finalize() {}

	virtual bool isCudaPass() const { return true; }
	virtual std::string getName() const { return "RenderPass"; }
	virtual bool enabled() const { return mEnable; }

protected:
	DeviceManager *getDeviceManager() const { return mDeviceManager; }
	size_t getFrameIndex() const;
	Vector2i getFrameSize() const;

	bool mEnable			   = true;
	Scene::SharedPtr mScene	   = nullptr;
	DeviceManager *mDeviceManager = nullptr;
};

class RenderPassFactory {
public:
	using map_type			  = std::map<std::string, std::function<RenderPass::SharedPtr(void)>>;
	using configured_map_type = std::map<std::string, std::function<RenderPass::SharedPtr(const json &)>>;

	template <typename T> static RenderPass::SharedPtr create() { return std::make_shared<T>(); }
	template <typename T> static RenderPass::SharedPtr deserialize(const json &serde) {
		auto ret = std::make_shared<T>();
		from_json(serde, *ret);
		return ret;
	}
	// nullptr (and a message on stderr) for an unregistered name, as the reference.
	static RenderPass::SharedPtr createInstance(const std::string &name);
	static RenderPass::SharedPtr deserizeInstance(const std::string &name, const json &serde);
	static std::vector<std::string> registeredNames();

protected:
	static map_type &getMap();
	static configured_map_type &getConfiguredMap();
};

template <typename T> class RenderPassRegister : RenderPassFactory {
public:
	explicit RenderPassRegister(const std::string &name) {
		getMap()[name]			 = &RenderPassFactory::create<T>;
		getConfiguredMap()[name] = &RenderPassFactory::deserialize<T>;
	}
};

#define KRR_REGISTER_PASS_DEC(name) static ::krr::RenderPassRegister<name> reg;
#define KRR_REGISTER_PASS_DEF(name) ::krr::RenderPassRegister<name> name::reg(#name);

} // namespace krr
