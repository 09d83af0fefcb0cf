// render_app.cpp - headless host loop (render_app.h).
#include "render_app.h"

#include <cstdio>

namespace krr {

RenderApp::RenderApp() : mContext(new RenderContext()) {}

RenderApp::~RenderApp() {
	try {
		finalize();
	} catch (...) {
	}
}

// renderer.cpp:258-315 ("resolution", "passes", "scene").
void RenderApp::loadConfig(const json &config) {
	try {
		if (config.contains("passes")) {
			for (const json &p : config.at("passes").items()) {
				const std::string name = p.at("name").asString();
				RenderPass::SharedPtr pass =
					p.contains("params") ? RenderPassFactory::deserizeInstance(name, p.at("params")) : RenderPassFactory::createInstance(name);
				if (!pass) {
					std::fprintf(stderr, "[krr] skipping render pass %s (not part of this library)\n", name.c_str());
					continue;
				}
				pass->setEnable(p.value("enable", true));
				addRenderPassToBack(pass);
			}
		}
		if (config.contains("scene")) setScene(Scene::fromJson(config.at("scene").dump()));
		if (config.contains("resolution")) {
			const json &r = config.at("resolution");
			if (!r.isArray() || r.size() != 2) throw std::runtime_error("\"resolution\" must be [width, height]");
			resize({(int) r[0].asNumber(), (int) r[1].asNumber()});
		} else if (mScene && mScene->getResolution() != Vector2i(0, 0)) {
			resize(mScene->getResolution());
		}
	} catch (const Error &) {
		throw;
	} catch (const std::exception &e) {
		throw Error(KRR_ERR_BAD_CONFIG, std::string("RenderApp::loadConfig: ") + e.what());
	}
}

void RenderApp::setScene(Scene::SharedPtr scene) {
	mScene = scene;
	if (mScene && mSize[0] > 0 && mSize[1] > 0) mScene->setAspectRatio((float) mSize[0] / (float) mSize[1]);
	mContext->setScene(scene);
	if (mInitialized)
		for (auto &p : mRenderPasses) p->setScene(scene);
}

void RenderApp::addRenderPassToBack(RenderPass::SharedPtr pass) {
	if (!pass) throw Error(KRR_ERR_INVALID_ARG, "RenderApp::addRenderPassToBack: null pass");
	pass->setDeviceManager(this);
	if (mSize != Vector2i(0, 0)) pass->resize(mSize);
	if (mInitialized) {
		pass->setScene(mScene);
		pass->initialize();
	}
	mRenderPasses.push_back(pass);
}

// Back buffer resize -> context (window.cpp:407-409) and every pass;
// camera aspect ratio (renderer.cpp:25-31).
void RenderApp::resize(const Vector2i size) {
	mSize = size;
	if (mScene && size[0] > 0 && size[1] > 0) mScene->setAspectRatio((float) size[0] / (float) size[1]);
	mContext->resize(size);
	for (auto &p : mRenderPasses) p->resize(size);
}

// renderer.cpp:324-331: the scene is set before the passes initialise.
void RenderApp::initialize() {
	for (auto &p : mRenderPasses) p->setScene(mScene);
	for (auto &p : mRenderPasses) p->initialize();
	mInitialized = true;
}

void RenderApp::finalize() {
	if (!mInitialized) return;
	for (auto &p : mRenderPasses) p->finalize();
	mInitialized = false;
}

void RenderApp::tick(float elapsedSeconds) {
	for (auto &p : mRenderPasses) p->tick(elapsedSeconds);
}

// window.cpp:412-425 + renderer.cpp:84-97, one iteration per frame.
void RenderApp::run(int frames) {
	if (!mInitialized) initialize();
	for (int f = 0; f < frames; f++) {
		mFrameIndex++; // 1-based (window.cpp:457)
		tick(0.f);
		for (auto &p : mRenderPasses) p->beginFrame(mContext.get());
		for (auto &p : mRenderPasses)
			if (p->enabled()) p->render(mContext.get());
		for (auto &p : mRenderPasses) p->endFrame(mContext.get());
	}
}

} // namespace krr
