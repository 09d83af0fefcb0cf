// wavefront_pass.h - the WavefrontPathTracer render pass on the gfx950
// library (src/render/wavefront/integrator.h:24-104 counterpart).
//
// Same name, JSON params and hooks as the reference pass; every hook
// forwards to the C ABI of include/krr_wavefront.h.  The params of
// integrator.h:96-103 keep their keys and defaults; the library's own keys
// (spp, rr_min_depth, ...) pass through unchanged.
#pragma once

#include "renderpass.h"

namespace krr {

class WavefrontPathTracer : public RenderPass {
public:
	using SharedPtr = std::shared_ptr<WavefrontPathTracer>;
	KRR_REGISTER_PASS_DEC(WavefrontPathTracer);

	WavefrontPathTracer() = default;
	~WavefrontPathTracer() override;
	WavefrontPathTracer(const WavefrontPathTracer &)			= delete;
	WavefrontPathTracer &operator=(const WavefrontPathTracer &) = delete;

	void resize(const Vector2i &size) override;
	void setScene(Scene::SharedPtr scene) override;
	void beginFrame(RenderContext *context) override;
	void render(RenderContext *context) override;
	void endFrame(RenderContext *context) override;
	void initialize() override;
	void finalize() override;
	std::string getName() const override { return "WavefrontPathTracer"; }

	// Counters of the last render (rays, node visits, kernel times).
	krr_stats getStats();
	// The library-side params after validation (to_json of the live pass).
	json getParams() const;
	krr_wf *handle() const { return mWf; }

	// integrator.h:76-83 members (same defaults)
	bool enableNEE	 = true;
	bool enableMedium = true;
	int maxDepth	 = 10;
	float probRR	 = 0.8f;
	bool enableClamp = false;
	float clampMax	 = 1e3f;
	// every other key of the params object (spp, rr_min_depth, ...)
	json extraParams = json::object();

	friend void to_json(json &j, const WavefrontPathTracer &p);
	friend void from_json(const json &j, WavefrontPathTracer &p);

private:
	void ensureCreated();
	void uploadSceneIfNeeded();
	krr_wf *mWf		 = nullptr;
	Vector2i mSize	 = {0, 0};
	bool mSceneDirty = false;
	bool mSized		 = false;
};

} // namespace krr
