// Parallel host memcpy for the host-buffer entry point: a few persistent threads copy the
// caller's (pageable) ray and record arrays to and from pinned staging buffers, which the DMA
// engines read and write at full PCIe rate. One pool per context; not thread-safe (a context
// has one caller at a time).
#pragma once
#include <sched.h>

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace prt {

class CopyPool {
  public:
	explicit CopyPool(int threads) : n_(std::max(1, threads)) {
		try {
			for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { loop(i); });
		} catch (...) {
			n_ = 1 + static_cast<int>(th_.size());  // work with the threads that started
		}
	}
	~CopyPool() {
		{
			std::lock_guard<std::mutex> l(m_);
			stop_ = true;
			++gen_;
		}
		cv_.notify_all();
		for (auto &t : th_) t.join();
	}
	CopyPool(const CopyPool &) = delete;
	CopyPool &operator=(const CopyPool &) = delete;

	// Copies `bytes` from src to dst, split across the pool; returns when all parts are done.
	void copy(void *dst, const void *src, size_t bytes) {
		if (n_ == 1 || bytes < (size_t{4} << 20)) {
			std::memcpy(dst, src, bytes);
			return;
		}
		{
			std::lock_guard<std::mutex> l(m_);
			dst_ = static_cast<char *>(dst);
			src_ = static_cast<const char *>(src);
			bytes_ = bytes;
			pending_ = n_ - 1;
			++gen_;
		}
		cv_.notify_all();
		part(0);
		std::unique_lock<std::mutex> l(m_);
		done_.wait(l, [this] { return pending_ == 0; });
	}

	// Threads for this process: three quarters of the CPUs it may run on, at most 12.
	static int default_threads() {
		cpu_set_t set;
		int cpus = 1;
		if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
		return std::max(1, std::min(12, cpus * 3 / 4));
	}

  private:
	// Part i of the current copy (64-byte aligned split points).
	void part(int i) {
		const size_t per = (bytes_ / n_ + 63) & ~static_cast<size_t>(63);
		const size_t b = std::min(bytes_, per * static_cast<size_t>(i));
		const size_t e = std::min(bytes_, b + per);
		if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
	}
	void loop(int i) {
		uint64_t seen = 0;
		for (;;) {
			{
				std::unique_lock<std::mutex> l(m_);
				cv_.wait(l, [&] { return gen_ != seen; });
				seen = gen_;
				if (stop_) return;
			}
			part(i);
			std::lock_guard<std::mutex> l(m_);
			if (--pending_ == 0) done_.notify_one();
		}
	}

	int n_;
	std::vector<std::thread> th_;
	std::mutex m_;
	std::condition_variable cv_, done_;
	uint64_t gen_ = 0;
	int pending_ = 0;
	bool stop_ = false;
	char *dst_ = nullptr;
	const char *src_ = nullptr;
	size_t bytes_ = 0;
};

}  // namespace prt
