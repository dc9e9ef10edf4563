/* Output streams with the muting switches the reference's tests use
 * (Streams::setStdMuteOut / setStdMuteErr, Lib/include/Utilities/TBTK/
 * Streams.h:67-163), reduced to what the Chebyshev path prints. */
#ifndef TBTK_AMD_STREAMS
#define TBTK_AMD_STREAMS

#include <string>

namespace TBTK {

class Streams {
public:
	static void setStdMuteOut();
	static void setStdMuteErr();
	static void out(const std::string &s);
	static void err(const std::string &s);
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
