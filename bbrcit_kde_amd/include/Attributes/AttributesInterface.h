// AttributesInterface.h -- an attribute type AttrT provides
// AttrT& merge(const AttrT&); merge(a, b) returns a copy of a merged with b.
// Same contract as the reference's include/Attributes/AttributesInterface.h.
#ifndef BBRCIT_AMD_ATTRIBUTESINTERFACE_H_
#define BBRCIT_AMD_ATTRIBUTESINTERFACE_H_

namespace bbrcit {

template <typename AttrT>
AttrT merge(const AttrT &lhs, const AttrT &rhs) {
  AttrT out = lhs;
  out.merge(rhs);
  return out;
}

}  // namespace bbrcit

#endif
