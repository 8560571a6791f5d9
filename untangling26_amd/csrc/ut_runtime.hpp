// Host-side runtime pieces shared by every subsystem: the context (device +
// stream + last error), owned device buffers and error propagation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdexcept>
#include <string>
#include <vector>

namespace ut
{

struct Error : std::runtime_error
{
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes returned through the C ABI.
enum Status : int
{
	kOk = 0,
	kCapacity = 1,	   // an output list was too small; counts say how large it must be
	kNotConverged = 2, // iterative solve hit its iteration budget
	kInvalid = -1,	   // bad argument
	kHipError = -2,	   // HIP runtime failure
	kInternal = -3,
};

#define UT_HIP_CHECK(expr)                                                                              \
	do                                                                                                  \
	{                                                                                                   \
		hipError_t _e = (expr);                                                                         \
		if (_e != hipSuccess)                                                                           \
			throw ::ut::Error(::ut::kHipError,                                                          \
				std::string(#expr) + " failed: " + hipGetErrorString(_e) + " at " + __FILE__ + ":"      \
					+ std::to_string(__LINE__));                                                        \
	} while (0)

#define UT_REQUIRE(cond, msg)                                                                           \
	do                                                                                                  \
	{                                                                                                   \
		if (!(cond))                                                                                    \
			throw ::ut::Error(::ut::kInvalid, std::string(msg));                                        \
	} while (0)

#define UT_LAUNCH_CHECK() UT_HIP_CHECK(hipGetLastError())

// A device allocation owned by the library.  Grows, never shrinks.
template <typename T>
struct DevBuf
{
	T*	   ptr = nullptr;
	size_t cap = 0;
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	~DevBuf() { release(); }
	void release()
	{
		if (ptr)
			(void)hipFree(ptr);
		ptr = nullptr;
		cap = 0;
	}
	void reserve(size_t n)
	{
		if (n <= cap && ptr)
			return;
		release();
		size_t bytes = (n ? n : 1) * sizeof(T);
		UT_HIP_CHECK(hipMalloc(&ptr, bytes));
		cap = n ? n : 1;
	}
	T* get() const { return ptr; }
};

struct Context
{
	int			device = 0;
	hipStream_t stream = nullptr;
	bool		own_stream = false;
	std::string last_error;

	void activate() const { UT_HIP_CHECK(hipSetDevice(device)); }
};

} // namespace ut
