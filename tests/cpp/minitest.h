// A minimal test harness for the C++ drop-in tests (gtest is not available
// here): TEST, EXPECT_*, and EXPECT_EXIT run in a forked child.
#pragma once

#include <sys/wait.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace minitest {
struct Case {
	const char *suite, *name;
	std::function<void()> fn;
};
inline std::vector<Case> &registry() {
	static std::vector<Case> r;
	return r;
}
inline int &failures() {
	static int f = 0;
	return f;
}
struct Registrar {
	Registrar(const char *s, const char *n, std::function<void()> f) {
		registry().push_back({s, n, f});
	}
};
inline void fail(const char *file, int line, const std::string &what) {
	fprintf(stderr, "%s:%d: FAILED %s\n", file, line, what.c_str());
	failures()++;
}
// Runs fn in a child process; returns its exit code (or -signal).
inline int run_child(const std::function<void()> &fn) {
	fflush(stdout);
	fflush(stderr);
	pid_t pid = fork();
	if (pid == 0) {
		fn();
		_exit(0);
	}
	int status = 0;
	waitpid(pid, &status, 0);
	if (WIFEXITED(status)) return WEXITSTATUS(status);
	return -WTERMSIG(status);
}
inline int run_all(const char *filter) {
	int ran = 0;
	for (auto &c : registry()) {
		std::string full = std::string(c.suite) + "." + c.name;
		if (filter && full.find(filter) == std::string::npos) continue;
		const int before = failures();
		c.fn();
		ran++;
		printf("[%s] %s\n", failures() == before ? "  OK  " : " FAIL ", full.c_str());
	}
	printf("%d tests, %d failures\n", ran, failures());
	return failures() ? 1 : 0;
}
}  // namespace minitest

#define TEST(suite, name)                                                       \
	static void suite##_##name();                                           \
	static minitest::Registrar reg_##suite##_##name(#suite, #name, suite##_##name); \
	static void suite##_##name()

#define EXPECT_TRUE(c)                                                          \
	do {                                                                    \
		if (!(c)) minitest::fail(__FILE__, __LINE__, #c);               \
	} while (0)
#define EXPECT_FALSE(c) EXPECT_TRUE(!(c))
#define EXPECT_EQ(a, b)                                                         \
	do {                                                                    \
		if (!((a) == (b))) minitest::fail(__FILE__, __LINE__, #a " == " #b); \
	} while (0)
#define EXPECT_DOUBLE_EQ(a, b) EXPECT_NEAR(a, b, 4 * 2.220446049250313e-16 * std::fabs((double)(b)))
#define EXPECT_NEAR(a, b, tol)                                                  \
	do {                                                                    \
		const double _a = (a), _b = (b);                                \
		if (!(std::fabs(_a - _b) <= (tol)))                             \
			minitest::fail(__FILE__, __LINE__,                      \
			               std::string(#a " ~ " #b ": ") + std::to_string(_a) + \
			                   " vs " + std::to_string(_b));         \
	} while (0)
#define ASSERT_EQ(a, b)                                                         \
	do {                                                                    \
		if (!((a) == (b))) {                                            \
			minitest::fail(__FILE__, __LINE__, #a " == " #b);       \
			return;                                                 \
		}                                                               \
	} while (0)
// The statement must exit the (forked) process with `code`.
#define EXPECT_EXIT_CODE(stmt, code)                                            \
	do {                                                                    \
		const int _rc = minitest::run_child([&]() { stmt; });           \
		if (_rc != (code))                                              \
			minitest::fail(__FILE__, __LINE__,                      \
			               std::string("exit code ") + std::to_string(_rc) + \
			                   " != " #code " for " #stmt);          \
	} while (0)

#define MINITEST_MAIN                                                           \
	int main(int argc, char **argv) {                                      \
		return minitest::run_all(argc > 1 ? argv[1] : nullptr);         \
	}
