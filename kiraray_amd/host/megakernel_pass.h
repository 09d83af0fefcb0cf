// megakernel_pass.h - the MegakernelPathTracer render pass on the gfx950
// library (src/render/megakernel/pathtracer.h:14-50 counterpart).
//
// Same name, JSON params (nee, max_depth, rr with the reference's defaults)
// and hooks as the reference pass; every hook forwards to the krr_mk_* entry
// points of include/krr_wavefront.h.  The library's own keys (spp,
// block_size, ...) pass through unchanged.
#pragma once

#include "renderpass.h"

namespace krr {

class MegakernelPathTracer : public RenderPass {
public:
	using SharedPtr = std::shared_ptr<MegakernelPathTracer>;
	KRR_REGISTER_PASS_DEC(MegakernelPathTracer);

	MegakernelPathTracer() = default;
	~MegakernelPathTracer() override;
	MegakernelPathTracer(const MegakernelPathTracer &)			  = delete;
	MegakernelPathTracer &operator=(const MegakernelPathTracer &) = delete;

	void resize(const Vector2i &size) override;
	void setScene(Scene::SharedPtr scene) override;
	void render(RenderContext *context) override;
	void initialize() override;
	void finalize() override;
	std::string getName() const override { return "MegakernelPathTracer"; }

	krr_stats getStats();
	json getParams() const;

	// LaunchParameters<MegakernelPathTracer> members set from JSON (path.h:22-27)
	bool enableNEE = true;
	int maxDepth   = 10;
	float probRR   = 0.8f;
	json extraParams = json::object();

	friend void to_json(json &j, const MegakernelPathTracer &p);
	friend void from_json(const json &j, MegakernelPathTracer &p);

private:
	void ensureCreated();
	void uploadSceneIfNeeded();
	krr_mk *mMk		 = nullptr;
	Vector2i mSize	 = {0, 0};
	bool mSceneDirty = false;
	bool mSized		 = false;
};

} // namespace krr
