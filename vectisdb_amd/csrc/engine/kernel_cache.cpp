// kernel_cache.cpp -- SHA-256 keyed LRU of compiled kernels.
//
// Behaviour follows cuda_engine/src/kernel_cache.cpp:149-222 (ownership through a registered deleter;
// replacement, eviction, decline and destruction all free the dropped handle). SHA-256 is written from
// FIPS 180-4: the round constants are the first 32 bits of the fractional parts of the cube roots of the
// first 64 primes and the initial state those of the square roots of the first 8 primes; both are
// computed here at start-up rather than tabulated.
#include "../../../include/kernel_cache.hpp"

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>

namespace vector_gpu {

namespace {

//! floor(frac(x) * 2^32) for x = p^(1/k), computed exactly enough for the FIPS constants: a long double
//! root is refined with integer Newton steps on the 32-bit fraction.
uint32_t FracRootBits(uint32_t prime, int k) {
	// Find the integer r with r^k <= prime * 2^(32k) < (r+1)^k; the low 32 bits of r are the fraction.
	long double approx = std::pow((long double)prime, 1.0L / k) * 4294967296.0L;
	unsigned __int128 r = (unsigned __int128)approx;
	unsigned __int128 target = (unsigned __int128)prime;
	for (int i = 0; i < k; i++) {
		target <<= 32;
	}
	auto power = [&](unsigned __int128 v) {
		// v^k may exceed 128 bits only when v is wildly off; k <= 3 and v < 2^40 keeps it in range.
		unsigned __int128 p = 1;
		for (int i = 0; i < k; i++) {
			p *= v;
		}
		return p;
	};
	while (power(r) > target) {
		r--;
	}
	while (power(r + 1) <= target) {
		r++;
	}
	return (uint32_t)(r & 0xffffffffu);
}

struct ShaConstants {
	std::array<uint32_t, 64> k;
	std::array<uint32_t, 8> h0;
	ShaConstants() {
		uint32_t primes[64];
		int found = 0;
		for (uint32_t c = 2; found < 64; c++) {
			bool is_prime = true;
			for (uint32_t d = 2; d * d <= c; d++) {
				if (c % d == 0) {
					is_prime = false;
					break;
				}
			}
			if (is_prime) {
				primes[found++] = c;
			}
		}
		for (int i = 0; i < 64; i++) {
			k[i] = FracRootBits(primes[i], 3);
		}
		for (int i = 0; i < 8; i++) {
			h0[i] = FracRootBits(primes[i], 2);
		}
	}
};

const ShaConstants &Constants() {
	static const ShaConstants c;
	return c;
}

inline uint32_t Rotr(uint32_t x, int n) {
	return (x >> n) | (x << (32 - n));
}

void Compress(std::array<uint32_t, 8> &h, const uint8_t *block) {
	const auto &K = Constants().k;
	uint32_t w[64];
	for (int i = 0; i < 16; i++) {
		w[i] = (uint32_t)block[4 * i] << 24 | (uint32_t)block[4 * i + 1] << 16 | (uint32_t)block[4 * i + 2] << 8 |
		       (uint32_t)block[4 * i + 3];
	}
	for (int i = 16; i < 64; i++) {
		uint32_t s0 = Rotr(w[i - 15], 7) ^ Rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
		uint32_t s1 = Rotr(w[i - 2], 17) ^ Rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
		w[i] = w[i - 16] + s0 + w[i - 7] + s1;
	}
	uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
	for (int i = 0; i < 64; i++) {
		uint32_t S1 = Rotr(e, 6) ^ Rotr(e, 11) ^ Rotr(e, 25);
		uint32_t ch = (e & f) ^ (~e & g);
		uint32_t t1 = hh + S1 + ch + K[i] + w[i];
		uint32_t S0 = Rotr(a, 2) ^ Rotr(a, 13) ^ Rotr(a, 22);
		uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
		uint32_t t2 = S0 + maj;
		hh = g;
		g = f;
		f = e;
		e = d + t1;
		d = c;
		c = b;
		b = a;
		a = t1 + t2;
	}
	h[0] += a;
	h[1] += b;
	h[2] += c;
	h[3] += d;
	h[4] += e;
	h[5] += f;
	h[6] += g;
	h[7] += hh;
}

std::string Sha256Hex(const std::string &data) {
	std::array<uint32_t, 8> h = Constants().h0;
	const uint8_t *p = reinterpret_cast<const uint8_t *>(data.data());
	const size_t n = data.size();
	size_t full = n / 64;
	for (size_t i = 0; i < full; i++) {
		Compress(h, p + 64 * i);
	}
	// Padding: 0x80, zeros, 64-bit big-endian bit length.
	uint8_t tail[128] = {0};
	size_t rem = n - full * 64;
	for (size_t i = 0; i < rem; i++) {
		tail[i] = p[full * 64 + i];
	}
	tail[rem] = 0x80;
	size_t tail_len = rem + 1 + 8 <= 64 ? 64 : 128;
	uint64_t bits = (uint64_t)n * 8;
	for (int i = 0; i < 8; i++) {
		tail[tail_len - 1 - i] = (uint8_t)(bits >> (8 * i));
	}
	Compress(h, tail);
	if (tail_len == 128) {
		Compress(h, tail + 64);
	}
	static const char *hex = "0123456789abcdef";
	std::string out(64, '0');
	for (int i = 0; i < 8; i++) {
		for (int j = 0; j < 4; j++) {
			uint8_t byte = (uint8_t)(h[i] >> (24 - 8 * j));
			out[8 * i + 2 * j] = hex[byte >> 4];
			out[8 * i + 2 * j + 1] = hex[byte & 15];
		}
	}
	return out;
}

} // namespace

KernelCache::~KernelCache() {
	for (auto &entry : entries_) {
		DeleteHandleLocked(entry.second.first);
	}
}

void KernelCache::SetHandleDeleter(HandleDeleter deleter) {
	std::lock_guard<std::mutex> guard(mutex_);
	if (deleter_ != nullptr && deleter_ != deleter) {
		throw std::logic_error("KernelCache::SetHandleDeleter: another deleter is already registered");
	}
	deleter_ = deleter;
}

void KernelCache::DeleteHandleLocked(CompiledKernelHandle handle) {
	if (deleter_ != nullptr && handle != nullptr) {
		deleter_(handle);
	}
}

CompiledKernelHandle KernelCache::Get(const std::string &source_hash) {
	std::lock_guard<std::mutex> guard(mutex_);
	auto it = entries_.find(source_hash);
	if (it == entries_.end()) {
		return nullptr;
	}
	lru_order_.splice(lru_order_.begin(), lru_order_, it->second.second);
	return it->second.first;
}

void KernelCache::Insert(const std::string &source_hash, CompiledKernelHandle handle) {
	std::lock_guard<std::mutex> guard(mutex_);
	if (max_entries_ == 0) {
		DeleteHandleLocked(handle);
		return;
	}
	auto it = entries_.find(source_hash);
	if (it != entries_.end()) {
		if (it->second.first != handle) {
			DeleteHandleLocked(it->second.first);
		}
		it->second.first = handle;
		lru_order_.splice(lru_order_.begin(), lru_order_, it->second.second);
		return;
	}
	if (entries_.size() >= max_entries_) {
		const std::string victim = lru_order_.back();
		lru_order_.pop_back();
		auto v = entries_.find(victim);
		if (v != entries_.end()) {
			DeleteHandleLocked(v->second.first);
			entries_.erase(v);
		}
	}
	lru_order_.push_front(source_hash);
	entries_.emplace(source_hash, std::make_pair(handle, lru_order_.begin()));
}

std::string KernelCache::HashSource(const std::string &cuda_source) {
	return Sha256Hex(cuda_source);
}

} // namespace vector_gpu
