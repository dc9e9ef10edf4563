/* Output streams with the muting switches the reference's tests use
 * (Lib/include/Utilities/TBTK/Streams.h:67-163): Streams::out and
 * Streams::err are std::ostreams writing to stdout and stderr until
 * setStdMuteOut / setStdMuteErr is called (no log file). */
#ifndef TBTK_AMD_STREAMS
#define TBTK_AMD_STREAMS

#include <ostream>
#include <string>

namespace TBTK {

class Streams {
public:
	static std::ostream out;
	static std::ostream err;
	static void setStdMuteOut();
	static void setStdMuteErr();
};

/* Global verbosity of Communicator (Lib/include/Utilities/TBTK/
 * Communicator.h): an object prints only if both are on. */
class Communicator {
public:
	explicit Communicator(bool verbose = false) : verbose(verbose) {}
	void setVerbose(bool v) { verbose = v; }
	bool getVerbose() const { return verbose; }
	static void setGlobalVerbose(bool v);
	static bool getGlobalVerbose();

private:
	bool verbose;
};

}  // namespace TBTK
#endif
