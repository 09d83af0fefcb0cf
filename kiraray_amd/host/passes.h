// passes.h - AccumulatePass and ErrorMeasurePass of the host layer
// (src/render/passes/accumulate/accumulate.h, errormeasure/errormeasure.h
// counterparts), so that a reference config's pass list
// [WavefrontPathTracer, AccumulatePass, ErrorMeasurePass] runs as it is.
// Both work in place on the RenderContext's colour texture, on its stream,
// through the C ABI (krr_accumulate, krr_error_metric).
#pragma once

#include <chrono>

#include "renderpass.h"

namespace krr {

// Running mean of the frames (accumulate.cu:30-79): "spp" frames at most
// (0 = no limit) in sum mode, or a moving average; float or double
// accumulation.  The colour texture is rewritten with the mean, so the
// passes after it see the accumulated image.
class AccumulatePass : public RenderPass {
public:
	using SharedPtr = std::shared_ptr<AccumulatePass>;
	KRR_REGISTER_PASS_DEC(AccumulatePass);
	enum class Mode { Accumulate = 0, MovingAverage = 1 };
	enum class Precision { Float = 0, Double = 1 };

	AccumulatePass() = default;
	~AccumulatePass() override;
	AccumulatePass(const AccumulatePass &)			  = delete;
	AccumulatePass &operator=(const AccumulatePass &) = delete;

	void resize(const Vector2i &size) override;
	void setScene(Scene::SharedPtr scene) override;
	void render(RenderContext *context) override;
	void finalize() override;
	std::string getName() const override { return "AccumulatePass"; }

	size_t getAccumCount() const { return mAccumCount; }
	void reset();

	friend void to_json(json &j, const AccumulatePass &p);
	friend void from_json(const json &j, AccumulatePass &p);

private:
	size_t mMaxAccumCount = 0, mAccumCount = 0;
	Mode mMode			  = Mode::Accumulate;
	Precision mPrecision  = Precision::Float;
	Vector2i mSize		  = {0, 0};
	void *mAccum		  = nullptr; // W*H float4 or double4
	size_t mAccumBytes	  = 0;
};

// Error of the colour texture against a reference image (errormeasure.cpp):
// metric 0 MSE, 1 MAPE, 2 SMAPE, 3 RelMSE (metrics.cu:80-128); evaluated
// every "interval" frames when "continuous", the series (timesteps,
// timepoints in seconds, data) written to "output" as JSON at finalize when
// "save".  "reference" is a float32 .npy of H x W x 4 (or x 3) in the
// colour texture's layout (row H-1-y holds image row y); the reference
// reads .exr through its image library, which is not part of this build.
class ErrorMeasurePass : public RenderPass {
public:
	using SharedPtr = std::shared_ptr<ErrorMeasurePass>;
	KRR_REGISTER_PASS_DEC(ErrorMeasurePass);

	ErrorMeasurePass() = default;
	~ErrorMeasurePass() override;
	ErrorMeasurePass(const ErrorMeasurePass &)			  = delete;
	ErrorMeasurePass &operator=(const ErrorMeasurePass &) = delete;

	void beginFrame(RenderContext *context) override;
	void render(RenderContext *context) override;
	void finalize() override;
	std::string getName() const override { return "ErrorMeasurePass"; }

	bool loadReferenceImage(const std::string &path);
	// {"timesteps": [...], "timepoints": [...], "data": [{name: value}]}
	json getResults() const;
	json getLastResult() const { return mLastResult; }
	void requestEvaluation() { mNeedsEvaluate = true; }

	friend void to_json(json &j, const ErrorMeasurePass &p);
	friend void from_json(const json &j, ErrorMeasurePass &p);

private:
	void reset();
	int mMetric = 3;
	std::string mReferencePath, mOutputPath = "error/result.json";
	bool mNeedsEvaluate = false, mContinuousEvaluate = false, mLogResults = false, mSaveResults = false;
	size_t mFrameNumber = 0, mEvaluateInterval = 1;
	Vector2i mReferenceSize = {0, 0};
	void *mReference		= nullptr; // W*H float4 on the device
	json mLastResult		= json::object();
	struct Evaluation {
		size_t timestep;
		double timepoint;
		json metrics;
	};
	std::vector<Evaluation> mResults;
	std::chrono::steady_clock::time_point mStart = std::chrono::steady_clock::now();
};

} // namespace krr
