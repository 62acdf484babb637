"""solaris_amd -- MI355X-native force-evaluation and integration path of
the Solaris N-body simulator, behind the C ABI in include/solaris_amd.h."""
