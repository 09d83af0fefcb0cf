// passes.cpp - AccumulatePass and ErrorMeasurePass over the C ABI.
#include "passes.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sys/stat.h>

namespace krr {

KRR_REGISTER_PASS_DEF(AccumulatePass);
KRR_REGISTER_PASS_DEF(ErrorMeasurePass);

namespace {

void checkHipCall(hipError_t e, const char *call) {
	if (e != hipSuccess) throw Error(KRR_ERR_HIP, std::string(call) + ": " + hipGetErrorString(e));
}
#define KRR_HIP_CALL(expr) checkHipCall((expr), #expr)

const char *const kMetricNames[] = {"MSE", "MAPE", "SMAPE", "RelMSE"}; // errormeasure.cpp:7

// An enum param as the reference writes it (NLOHMANN_JSON_SERIALIZE_ENUM
// strings, accumulate.h:69-77, metrics.h:9-14; an unknown string maps to the
// first value, as there) or as its integer.
int enumParam(const json &j, const char *key, const std::vector<std::string> &names, int def) {
	if (!j.isObject() || !j.contains(key)) return def;
	const json &v = j.at(key);
	if (v.isString()) {
		const std::string s = v.asString();
		for (size_t i = 0; i < names.size(); i++)
			if (names[i] == s) return (int) i;
		return 0;
	}
	const double d = v.asNumber();
	if (d != (double) (int) d || d < 0 || d >= (double) names.size())
		throw std::runtime_error(std::string(key) + " out of range");
	return (int) d;
}

const std::vector<std::string> kModeNames	   = {"accumulate", "moving average"};
const std::vector<std::string> kPrecisionNames = {"float", "double"};
const std::vector<std::string> kMetricKeys	   = {"mse", "mape", "smape", "rel_mse"};

// Reads a little-endian float32 C-order .npy of shape (H, W, 3) or (H, W, 4)
// into RGBA (alpha 1 for three channels).
std::vector<float> readNpyRgba(const std::string &path, int &H, int &W) {
	std::ifstream f(path, std::ios::binary);
	if (!f) throw Error(KRR_ERR_INVALID_ARG, "cannot open " + path);
	char magic[8];
	f.read(magic, 8);
	if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw Error(KRR_ERR_INVALID_ARG, path + ": not a .npy file");
	uint32_t hlen = 0;
	if (magic[6] == 1) {
		uint16_t h16 = 0;
		f.read(reinterpret_cast<char *>(&h16), 2);
		hlen = h16;
	} else f.read(reinterpret_cast<char *>(&hlen), 4);
	std::string header(hlen, ' ');
	f.read(&header[0], hlen);
	if (!f) throw Error(KRR_ERR_INVALID_ARG, path + ": truncated header");
	if (header.find("'<f4'") == std::string::npos) throw Error(KRR_ERR_INVALID_ARG, path + ": dtype must be little-endian float32");
	if (header.find("'fortran_order': False") == std::string::npos) throw Error(KRR_ERR_INVALID_ARG, path + ": must be C order");
	const size_t s = header.find("'shape': (");
	if (s == std::string::npos) throw Error(KRR_ERR_INVALID_ARG, path + ": no shape");
	int dims[3] = {0, 0, 0}, nd = 0;
	size_t p = s + 10;
	while (nd < 3 && p < header.size() && header[p] != ')') {
		while (p < header.size() && (header[p] == ' ' || header[p] == ',')) p++;
		if (p >= header.size() || header[p] == ')') break;
		dims[nd++] = std::atoi(header.c_str() + p);
		while (p < header.size() && header[p] >= '0' && header[p] <= '9') p++;
	}
	if (nd != 3 || (dims[2] != 3 && dims[2] != 4) || dims[0] < 1 || dims[1] < 1)
		throw Error(KRR_ERR_INVALID_ARG, path + ": shape must be (H, W, 3) or (H, W, 4)");
	H = dims[0];
	W = dims[1];
	const size_t n = (size_t) H * W;
	std::vector<float> src(n * dims[2]);
	f.read(reinterpret_cast<char *>(src.data()), (std::streamsize) (src.size() * sizeof(float)));
	if (!f) throw Error(KRR_ERR_INVALID_ARG, path + ": truncated data");
	std::vector<float> rgba(n * 4, 1.f);
	for (size_t i = 0; i < n; i++)
		for (int c = 0; c < dims[2]; c++) rgba[4 * i + c] = src[dims[2] * i + c];
	return rgba;
}

void makeParentDirs(const std::string &path) {
	for (size_t i = 1; i < path.size(); i++)
		if (path[i] == '/') (void) ::mkdir(path.substr(0, i).c_str(), 0755);
}

} // namespace

// ------------------------------------------------------- AccumulatePass

// accumulate.h:31-49 (keys and defaults; the enums as the reference's strings)
void to_json(json &j, const AccumulatePass &p) {
	j = json::object();
	j.set("spp", json((double) p.mMaxAccumCount));
	j.set("mode", json(kModeNames[(int) p.mMode]));
	j.set("precision", json(kPrecisionNames[(int) p.mPrecision]));
}

void from_json(const json &j, AccumulatePass &p) {
	try {
		if (!j.isObject() && !j.isNull()) throw std::runtime_error("params must be a JSON object");
		const double spp = j.value("spp", 0.0);
		const int mode	 = enumParam(j, "mode", kModeNames, 0);
		const int prec	 = enumParam(j, "precision", kPrecisionNames, 0);
		if (!(spp >= 0)) throw std::runtime_error("spp must be >= 0");
		p.mMaxAccumCount = (size_t) spp;
		p.mMode			 = (AccumulatePass::Mode) mode;
		p.mPrecision	 = (AccumulatePass::Precision) prec;
	} catch (const Error &) {
		throw;
	} catch (const std::exception &e) {
		throw Error(KRR_ERR_BAD_CONFIG, std::string("AccumulatePass params: ") + e.what());
	}
}

AccumulatePass::~AccumulatePass() {
	if (mAccum) (void) hipFree(mAccum);
}

// accumulate.cu:24-29; the buffer follows the frame size and precision and
// is (re)allocated at the next render, so that a pass can be configured
// without a device
void AccumulatePass::reset() { mAccumCount = 0; }

void AccumulatePass::resize(const Vector2i &size) {
	mSize = size;
	reset();
}

void AccumulatePass::setScene(Scene::SharedPtr scene) {
	RenderPass::setScene(scene);
	if (mSize[0] > 0) reset(); // a new scene starts a new mean (accumulate.cu:55)
}

// accumulate.cu:54-77
void AccumulatePass::render(RenderContext *context) {
	if (!mEnable) return;
	const CudaRenderTarget t = context->getColorTexture()->getCudaRenderTarget();
	if (t.width != mSize[0] || t.height != mSize[1]) throw Error(KRR_ERR_INVALID_ARG, "AccumulatePass::render: colour texture size differs");
	const size_t bytes = (size_t) mSize[0] * mSize[1] * (mPrecision == Precision::Double ? 32 : 16);
	if (bytes != mAccumBytes || !mAccum) {
		if (mAccum) KRR_HIP_CALL(hipFree(mAccum));
		mAccum		= nullptr;
		mAccumBytes = 0;
		KRR_HIP_CALL(hipMalloc(&mAccum, bytes));
		mAccumBytes = bytes;
		mAccumCount = 0;
	}
	KRR_CALL(krr_accumulate(mAccum, t.data, (int64_t) t.width * t.height, (int64_t) mAccumCount, (int64_t) mMaxAccumCount, (int32_t) mMode,
							mPrecision == Precision::Double ? 64 : 32, context->getCudaStream()));
	if (!mMaxAccumCount || mAccumCount < mMaxAccumCount) mAccumCount++;
}

void AccumulatePass::finalize() {
	if (mAccum) KRR_HIP_CALL(hipFree(mAccum));
	mAccum		= nullptr;
	mAccumBytes = 0;
}

// ----------------------------------------------------- ErrorMeasurePass

// errormeasure.h:52-73 (keys and defaults); "output" names the JSON file
// the series is saved to (the reference derives it from its output folder).
void to_json(json &j, const ErrorMeasurePass &p) {
	j = json::object();
	j.set("metric", json(kMetricKeys[p.mMetric]));
	j.set("reference", json(p.mReferencePath));
	j.set("continuous", json(p.mContinuousEvaluate));
	j.set("interval", json((double) p.mEvaluateInterval));
	j.set("log", json(p.mLogResults));
	j.set("save", json(p.mSaveResults));
	j.set("output", json(p.mOutputPath));
}

void from_json(const json &j, ErrorMeasurePass &p) {
	try {
		if (!j.isObject() && !j.isNull()) throw std::runtime_error("params must be a JSON object");
		p.mMetric			  = enumParam(j, "metric", kMetricKeys, 3);
		p.mContinuousEvaluate = j.value("continuous", false);
		const double interval = j.value("interval", 1.0);
		p.mLogResults		  = j.value("log", false);
		p.mSaveResults		  = j.value("save", false);
		p.mOutputPath		  = j.value("output", "error/result.json");
		if (!(interval >= 1)) throw std::runtime_error("interval must be >= 1");
		p.mEvaluateInterval = (size_t) interval;
	} catch (const Error &) {
		throw;
	} catch (const std::exception &e) {
		throw Error(KRR_ERR_BAD_CONFIG, std::string("ErrorMeasurePass params: ") + e.what());
	}
	if (j.isObject() && j.contains("reference")) p.loadReferenceImage(j.value("reference", ""));
}

ErrorMeasurePass::~ErrorMeasurePass() {
	if (mReference) (void) hipFree(mReference);
}

// errormeasure.cpp:90-114
bool ErrorMeasurePass::loadReferenceImage(const std::string &path) {
	int H = 0, W = 0;
	const std::vector<float> rgba = readNpyRgba(path, H, W);
	if (mReference) KRR_HIP_CALL(hipFree(mReference));
	mReference = nullptr;
	KRR_HIP_CALL(hipMalloc(&mReference, rgba.size() * sizeof(float)));
	KRR_HIP_CALL(hipMemcpy(mReference, rgba.data(), rgba.size() * sizeof(float), hipMemcpyHostToDevice));
	mReferenceSize = {W, H};
	mReferencePath = path;
	reset();
	return true;
}

// errormeasure.cpp:84-89
void ErrorMeasurePass::reset() {
	mFrameNumber   = 0;
	mNeedsEvaluate = false;
	mLastResult	   = json::object();
	mStart		   = std::chrono::steady_clock::now();
}

// errormeasure.cpp:11-15
void ErrorMeasurePass::beginFrame(RenderContext *context) {
	if (!mFrameNumber) reset();
	mFrameNumber++;
	mNeedsEvaluate |= mContinuousEvaluate && (mFrameNumber % mEvaluateInterval == 0);
}

// errormeasure.cpp:17-37
void ErrorMeasurePass::render(RenderContext *context) {
	if (!mEnable || !mNeedsEvaluate || !mReference) return;
	const CudaRenderTarget t = context->getColorTexture()->getCudaRenderTarget();
	if (t.width != mReferenceSize[0] || t.height != mReferenceSize[1])
		throw Error(KRR_ERR_INVALID_ARG, "ErrorMeasurePass: reference image size does not match frame size");
	double value = 0.0;
	KRR_CALL(krr_error_metric(mMetric, t.data, (const float *) mReference, (int64_t) t.width * t.height, &value, context->getCudaStream()));
	mLastResult = json::object();
	mLastResult.set(kMetricNames[mMetric], json(value));
	if (mLogResults) std::fprintf(stderr, "[krr] evaluating frame #%zu: %s\n", mFrameNumber, mLastResult.dump().c_str());
	if (mSaveResults)
		mResults.push_back({mFrameNumber, std::chrono::duration<double>(std::chrono::steady_clock::now() - mStart).count(), mLastResult});
	mNeedsEvaluate = false;
}

json ErrorMeasurePass::getResults() const {
	json steps = json::array(), points = json::array(), data = json::array();
	for (const Evaluation &e : mResults) {
		steps.push(json((double) e.timestep));
		points.push(json(e.timepoint));
		data.push(e.metrics);
	}
	json r = json::object();
	r.set("timesteps", steps);
	r.set("timepoints", points);
	r.set("data", data);
	return r;
}

// errormeasure.cpp:39-57
void ErrorMeasurePass::finalize() {
	if (mSaveResults) {
		makeParentDirs(mOutputPath);
		std::ofstream f(mOutputPath);
		if (!f) throw Error(KRR_ERR_INVALID_ARG, "ErrorMeasurePass: cannot write " + mOutputPath);
		f << getResults().dump();
		std::fprintf(stderr, "[krr] saved error evaluation data to %s\n", mOutputPath.c_str());
	}
	if (mReference) KRR_HIP_CALL(hipFree(mReference));
	mReference = nullptr;
}

} // namespace krr
