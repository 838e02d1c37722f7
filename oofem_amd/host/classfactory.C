#include "classfactory.h"

namespace oofem {

ClassFactory &GiveClassFactory()
{
    static ClassFactory cf;
    return cf;
}

}  // namespace oofem
