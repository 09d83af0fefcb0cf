// render_app.h - headless host loop over render passes (RenderApp,
// src/main/renderer.cpp:84-97,258-331, with the frame bookkeeping of
// DeviceManager, src/core/window.cpp:412-425,457,487-490).
//
// loadConfig understands the reference's config keys "resolution" and
// "passes" ([{name, enable, params}]); "scene" is the procedural scene
// JSON of krr_scene_create.  Each frame: tick(all) -> beginFrame(all) ->
// render(enabled) -> endFrame(all), with a 1-based frame index.
#pragma once

#include "renderpass.h"

namespace krr {

class RenderApp : public DeviceManager {
public:
	RenderApp();
	~RenderApp() override;

	// Passes named in "passes" that are not registered are skipped with a
	// message (the reference would dereference a null pass).
	void loadConfig(const json &config);
	void setScene(Scene::SharedPtr scene);
	Scene::SharedPtr getScene() const { return mScene; }
	void addRenderPassToBack(RenderPass::SharedPtr pass);
	const std::vector<RenderPass::SharedPtr> &getRenderPasses() const { return mRenderPasses; }
	// Resizes the context and every pass, and sets the scene camera's aspect
	// ratio to width / height (RenderApp::backBufferResized, renderer.cpp:25-31).
	void resize(const Vector2i size);

	// setScene then initialize on every pass (renderer.cpp:324-331).
	void initialize();
	void finalize();
	// Runs n frames of the host loop.
	void run(int frames);

	size_t getFrameIndex() const override { return mFrameIndex; }
	Vector2i getFrameSize() const override { return mSize; }
	RenderContext *getRenderContext() { return mContext.get(); }

private:
	void tick(float elapsedSeconds);
	std::unique_ptr<RenderContext> mContext;
	std::vector<RenderPass::SharedPtr> mRenderPasses;
	Scene::SharedPtr mScene;
	Vector2i mSize	   = {0, 0};
	size_t mFrameIndex = 0;
	bool mInitialized  = false;
};

} // namespace krr
